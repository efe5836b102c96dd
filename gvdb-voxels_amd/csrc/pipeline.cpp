// pipeline.cpp -- the stages of the point path over G grids at once.
//
// Each stage validates every grid, does the per-grid host work (pool sizes,
// buffer growth), then issues one launch per kernel over all grids
// (kernels.h, multi.h).  Host round trips happen once per stage for the
// whole set, not once per grid.  A single VolumeGVDB call is the same code
// with G = 1.  Grids of one call share the stream and staging in `Exec`.
#include "volume.h"
#include "graph.h"
#include "../../include/gvdb_amd.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace gvdb {

#define XCHK(v, x)                                                   \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) return (v)->hip_fail(e_, #x);          \
    } while (0)

// Point radius: the dilation marks every brick of [int(p) - r, int(p) + 2r]
// and the insert every subcell of [int(p) - r, int(p) + 1 + r] per point, one
// thread per point, so the radius bounds one thread's loop: at 1024 voxels
// that is 5.7e7 bricks or, with 4-voxel subcells, 1.3e8 subcells.  Totals
// beyond the tables' ranges (2^34 domain bricks, 2^31 entries) are refused
// where they are counted.
static inline bool radius_ok(float r) { return r >= 0.0f && r <= 1024.0f; }

static inline float key_float(uint32_t k) {
    uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static int drain(Volume* v, const Exec& ex, int G) {
    XCHK(v, hipMemcpyAsync(ex.hsmall, ex.dsmall, size_t(G) * SMALL_WORDS * 4, hipMemcpyDeviceToHost, ex.s));
    // recording: the copy is part of the graph; its words are read after a replay
    if (ex.cap) return GVDB_OK;
    XCHK(v, hipStreamSynchronize(ex.s));
    if (ex.ar) ex.ar->reset();
    return GVDB_OK;
}

#define FAILG(g, code)               \
    do {                             \
        if (failed) *failed = (g);   \
        return (code);               \
    } while (0)

static bool keyed_build_wanted(const BitDomain& d0);

// ------------------------------------------------------- RebuildTopology ----
int Volume::multi_rebuild_topology(Volume* const* vs, int G, const int* n, const float* radius, const float* orig3,
                                   const Exec& ex, int* failed) {
    if (ex.cap) FAILG(0, vs[0]->not_recordable("RebuildTopology"));
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (v->levels_ == 0) FAILG(g, v->fail(GVDB_ERR_STATE, "RebuildTopology before Configure"));
        if (n[g] <= 0 || !v->pts_.base || n[g] > v->pts_.n)
            FAILG(g, v->fail(GVDB_ERR_INVALID, "RebuildTopology: point count exceeds SetPoints"));
        if (!radius_ok(radius[g]))
            FAILG(g, v->fail(GVDB_ERR_UNSUPPORTED, "RebuildTopology: radius must be in [0, 1024]"));
    }
    Volume* v0 = vs[0];
    if (G == 1) v0->stage_begin(ST_TOPO);
    std::vector<BboxArgs> ba(G);
    for (int g = 0; g < G; g++) {
        PointSrc P = vs[g]->pts_;
        P.n = n[g];
        XCHK(vs[g], vs[g]->bbox_part_.ensure(size_t(BBOX_PART_WORDS) * 4));
        ba[g] = BboxArgs{P, {orig3[3 * g], orig3[3 * g + 1], orig3[3 * g + 2]}, ex.dsmall + g * SMALL_WORDS + SM_BBOX,
                         vs[g]->bbox_part_.as<uint32_t>()};
    }
    XCHK(v0, launch_bbox_multi(ba.data(), G, ex.ar, ex.s));
    int rc = drain(v0, ex, G);
    if (rc) return rc;

    // per grid: bounds, depth and the reference's choice between a full and
    // an incremental build (gvdb_volume_gvdb.cpp:1195-1217): full after
    // RequestFullRebuild or Configure (mRebuildTopo), or when the depth is not
    // mCurrDepth, which only that branch sets; incremental otherwise.  An
    // incremental call whose depth is not the root's level (a requested full
    // rebuild changed it since) rebuilds fully here, where the reference
    // would walk a tree of another depth.
    std::vector<float> mnmx(6 * size_t(G));
    std::vector<int> depth(G), root(3 * size_t(G)), full_idx;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        const uint32_t* hk = ex.hsmall + g * SMALL_WORDS + SM_BBOX;
        for (int a = 0; a < 3; a++) {
            mnmx[6 * g + a] = key_float(hk[a]);
            mnmx[6 * g + 3 + a] = key_float(hk[3 + a]);
            // GetBoundingBox padding (gvdb_volume_gvdb.cpp:5452-5460)
            v->posMin_[a] = mnmx[6 * g + a] - 1;
            v->posMax_[a] = mnmx[6 * g + 3 + a] + 2;
            v->posMin_[a] = 0 > v->posMin_[a] ? 0 : v->posMin_[a];
            v->posMax_[a] = 0 > v->posMax_[a] ? 0 : v->posMax_[a];
        }
        depth[g] = v->determine_depth(&root[3 * g]);
        if (depth[g] == 0) FAILG(g, v->fail(GVDB_ERR_DEPTH, "Points extents smaller than 1 brick. Not yet supported."));
        if (depth[g] < 0 || depth[g] > 5 || depth[g] >= v->levels_)
            FAILG(g, v->fail(GVDB_ERR_DEPTH, "Point extents exceed 5 levels."));
    }
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (v->rebuild_topo_) {
            v->clear();
        } else if (depth[g] != v->curr_depth_ || v->root_ == ID_UNDEFL || int((v->root_ >> 8) & 0xFF) != depth[g]) {
            v->clear();
            v->rebuild_topo_ = true;
            v->curr_depth_ = depth[g];
        }
        if (v->rebuild_topo_) {
            full_idx.push_back(g);
            continue;
        }
        const Exec eg{ex.s, ex.ar, ex.dsmall + g * SMALL_WORDS, ex.hsmall + g * SMALL_WORDS};
        rc = v->incremental_topology(n[g], radius[g], orig3 + 3 * g, depth[g], false, &mnmx[6 * g], eg);
        if (rc) FAILG(g, rc);
    }
    // level-0 boxes too large for bitmaps (or the keyed path forced): the
    // keyed build, one grid at a time
    {
        std::vector<int> bitmap_idx;
        for (int g : full_idx) {
            Volume* v = vs[g];
            BitDomain dom[MAXLEV];
            if (!plan_domains(v, depth[g], &root[3 * g], &mnmx[6 * g], &mnmx[6 * g + 3], radius[g], dom))
                FAILG(g, v->fail(GVDB_ERR_RANGE, "RebuildTopology: brick coordinates exceed the 16-bit keys of the reference"));
            if (!keyed_build_wanted(dom[0])) {
                bitmap_idx.push_back(g);
                continue;
            }
            const Exec eg{ex.s, ex.ar, ex.dsmall + g * SMALL_WORDS, ex.hsmall + g * SMALL_WORDS};
            rc = v->keyed_build(n[g], radius[g], orig3 + 3 * g, depth[g], &root[3 * g], dom, eg);
            if (rc) FAILG(g, rc);
        }
        full_idx.swap(bitmap_idx);
    }
    if (!full_idx.empty()) {
        const int F = int(full_idx.size());
        std::vector<Volume*> fv(F);
        std::vector<int> fn(F), fd(F), fr(3 * size_t(F));
        std::vector<float> frad(F), fo(3 * size_t(F)), fm(6 * size_t(F));
        for (int k = 0; k < F; k++) {
            const int g = full_idx[k];
            fv[k] = vs[g];
            fn[k] = n[g];
            fd[k] = depth[g];
            frad[k] = radius[g];
            for (int a = 0; a < 3; a++) { fr[3 * k + a] = root[3 * g + a]; fo[3 * k + a] = orig3[3 * g + a]; }
            for (int a = 0; a < 6; a++) fm[6 * k + a] = mnmx[6 * g + a];
        }
        int f = -1;
        rc = multi_full_build(fv.data(), F, fn.data(), frad.data(), fo.data(), fd.data(), fr.data(), fm.data(), ex, &f);
        if (rc) {
            if (failed && f >= 0) *failed = full_idx[f];
            return rc;
        }
    }
    if (G == 1) v0->stage_end(ST_TOPO);
    return GVDB_OK;
}

// The brick box of the points at each level below the root: the points'
// bounds widened by the dilation ([int(p) - r, int(p) + 2r]) and one voxel,
// clipped to the root node; level L holds the level-L cells of that box.
// False when level-0 coordinates exceed the reference's 16-bit keys.
bool Volume::plan_domains(const Volume* v, int D, const int rootPos[3], const float* mn, const float* mx, float radius,
                  BitDomain* dom) {
    const int rr = radius >= 1.0f ? int(ceilf(radius)) : 0;
    long long lo_v[3], hi_v[3];
    for (int a = 0; a < 3; a++) {
        long long re = (long long)rootPos[a] + v->range_[D] - 1;
        long long l = std::isfinite(mn[a]) ? (long long)floorf(std::max(mn[a], 0.0f)) - rr - 1 : 0;
        long long h = std::isfinite(mx[a]) ? (long long)floorf(std::max(mx[a], 0.0f)) + 2 * rr + 1 : 0;
        lo_v[a] = std::max<long long>(rootPos[a], l);
        hi_v[a] = std::min<long long>(re, std::max<long long>(h, lo_v[a]));
    }
    if (hi_v[0] / v->range_[0] > 65535 || hi_v[1] / v->range_[0] > 65535 || hi_v[2] / v->range_[0] > 65535) return false;
    for (int L = 0; L < D; L++) {
        int shift = 0;
        for (int k = 0; k <= L; k++) shift += v->logdim_[k];
        BitDomain& d = dom[L];
        d.nbits = 1;
        for (int a = 0; a < 3; a++) {
            d.lo[a] = int(lo_v[a] >> shift);
            d.dim[a] = int((hi_v[a] >> shift) - d.lo[a] + 1);
            d.nbits *= u64(d.dim[a]);
        }
        d.nwords = (d.nbits + 31) / 32;
    }
    return true;
}

// Level-0 boxes above this many bricks take the keyed build (bitmaps of
// 2^34 bits are 2 GiB per level and set).
static const u64 BITMAP_MAX_BITS = u64(1) << 34;

static bool keyed_build_wanted(const BitDomain& d0) { return d0.nbits > BITMAP_MAX_BITS || knob(KNOB_MARK_PATH) == 4; }

// The full build of RebuildTopology (ActivateBricksGPU + ActivateExtraBricksGPU)
// for grids already cleared, with their depth, root corner and point bounds.
int Volume::multi_full_build(Volume* const* vs, int G, const int* n, const float* radius, const float* orig3,
                             const int* depth, const int* rootPos, const float* mnmx, const Exec& ex, int* failed) {
    Volume* v0 = vs[0];
    int rc = 0;
    // per grid: depth, root, domains, scratch
    struct Plan {
        int D = 0;
        bool dilate = false;
        int rootPos[3] = {0, 0, 0};
        BitDomain dom[MAXLEV];
        u64 bm_off[MAXLEV], rk_off[MAXLEV], tmp_off[2 * MAXLEV];
        uint32_t* masks = nullptr;
        uint8_t* bytes = nullptr;
        bool btab = false;
    };
    std::vector<Plan> plan(G);
    std::vector<FillArgs> fills;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        Plan& pl = plan[g];
        const float* mn = mnmx + 6 * g;
        const float* mx = mnmx + 6 * g + 3;
        for (int a = 0; a < 3; a++) pl.rootPos[a] = rootPos[3 * g + a];
        pl.D = depth[g];
        pl.dilate = radius[g] >= 1.0f;
        const int D = pl.D;
        if (!plan_domains(v, D, pl.rootPos, mn, mx, radius[g], pl.dom))
            FAILG(g, v->fail(GVDB_ERR_RANGE, "RebuildTopology: brick coordinates exceed the 16-bit keys of the reference"));
        u64 bm_words = 0, rk_words = 0, tmp_words = 0;
        for (int L = 0; L < D; L++) {
            const BitDomain& d = pl.dom[L];
            pl.bm_off[L] = bm_words;
            bm_words += d.nwords * (pl.dilate ? 2 : 1);
            pl.rk_off[L] = rk_words;
            rk_words += d.nwords * (pl.dilate ? 2 : 1);
            for (int k = 0; k < (pl.dilate ? 2 : 1); k++) {
                pl.tmp_off[2 * L + k] = tmp_words;
                tmp_words += scan_tmp_words(d.nwords);
            }
        }
        if (pl.dom[0].nbits > BITMAP_MAX_BITS)
            FAILG(g, v->fail(GVDB_ERR_RANGE, "RebuildTopology: brick box of the points exceeds 2^34 bricks"));
        XCHK(v, v->bitmaps_.ensure(bm_words * 4 + 256));
        XCHK(v, v->ranks_.ensure(rk_words * 4 + 256));
        XCHK(v, v->scan_tmp_.ensure(tmp_words * 4 + 256));
        fills.push_back(FillArgs{v->bitmaps_.p, bm_words, 0, 0});
        // marks: a byte per domain brick (own-brick marks, then dilation
        // marks, whole 32-byte groups) up to 2^27 bricks; else direction masks
        // (4 bytes per domain brick) while that stays comparable to the point
        // stream; else bits set directly
        // (the mark-path knob forces one: every path marks the same bits)
        const int force = knob(KNOB_MARK_PATH);
        const bool bytes_ok = pl.dom[0].nbits <= (u64(1) << 27);
        const bool masks_ok = pl.dom[0].nbits <= (u64(1) << 30);
        int path = bytes_ok ? 1 : (pl.dom[0].nbits <= std::max(u64(1) << 22, u64(n[g]) * 4) ? 2 : 3);
        if (force == 1 && bytes_ok) path = 1;
        if (force == 2 && masks_ok) path = 2;
        if (force == 3) path = 3;
        if (path == 1) {
            const u64 nbytes = pl.dom[0].nwords * 32 * (pl.dilate ? 2 : 1);
            XCHK(v, v->masks_.ensure(size_t(nbytes)));
            pl.bytes = reinterpret_cast<uint8_t*>(v->masks_.p);
            fills.push_back(FillArgs{v->masks_.p, nbytes / 4, 0, 0});
        } else if (path == 2) {
            XCHK(v, v->masks_.ensure(size_t(pl.dom[0].nbits) * 4));
            pl.masks = v->masks_.as<uint32_t>();
            fills.push_back(FillArgs{pl.masks, pl.dom[0].nbits, 0, 0});
        }
    }
    XCHK(v0, launch_fill_multi(fills.data(), int(fills.size()), ex.ar, ex.s));
    auto A = [&](int g, int L) { return vs[g]->bitmaps_.as<uint32_t>() + plan[g].bm_off[L]; };
    auto B = [&](int g, int L) {
        return plan[g].dilate ? vs[g]->bitmaps_.as<uint32_t>() + plan[g].bm_off[L] + plan[g].dom[L].nwords : nullptr;
    };
    std::vector<MarkArgs> ma(G);
    int maxD = 0;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        PointSrc P = v->pts_;
        P.n = n[g];
        ma[g] = MarkArgs{P, {orig3[3 * g], orig3[3 * g + 1], orig3[3 * g + 2]}, radius[g], plan[g].dilate ? 1 : 0,
                         int(v->range_[0]), 1.0f / float(v->range_[0]), plan[g].dom[0], A(g, 0), B(g, 0), plan[g].masks,
                         plan[g].bytes};
        maxD = std::max(maxD, plan[g].D);
    }
    XCHK(v0, launch_mark_points_multi(ma.data(), G, ex.ar, ex.s));
    for (int L = 1; L < maxD; L++) {
        std::vector<ParentArgs> pa;
        for (int g = 0; g < G; g++)
            if (L < plan[g].D)
                pa.push_back(ParentArgs{plan[g].dom[L - 1], A(g, L - 1), B(g, L - 1), vs[g]->logdim_[L], plan[g].dom[L],
                                        A(g, L), B(g, L)});
        XCHK(v0, launch_mark_parents_multi(pa.data(), int(pa.size()), ex.ar, ex.s));
    }
    std::vector<ScanArgs> sa;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        uint32_t* rk = v->ranks_.as<uint32_t>();
        uint32_t* tmp = v->scan_tmp_.as<uint32_t>();
        uint32_t* counts = ex.dsmall + g * SMALL_WORDS + SM_COUNTS;
        for (int L = 0; L < plan[g].D; L++) {
            const BitDomain& d = plan[g].dom[L];
            sa.push_back(ScanArgs{A(g, L), nullptr, 1, rk + plan[g].rk_off[L], d.nwords, tmp + plan[g].tmp_off[2 * L],
                                  counts + 2 * L});
            if (plan[g].dilate)
                sa.push_back(ScanArgs{A(g, L), B(g, L), 2, rk + plan[g].rk_off[L] + d.nwords, d.nwords,
                                      tmp + plan[g].tmp_off[2 * L + 1], counts + 2 * L + 1});
        }
    }
    XCHK(v0, launch_scan_multi(sa.data(), int(sa.size()), ex.ar, ex.s));
    rc = drain(v0, ex, G);
    if (rc) return rc;

    // per grid: pools sized from the counts, then nodes and links in one launch per kernel
    std::vector<RootArgs> ra;
    std::vector<FillArgs> cf;
    std::vector<EmitArgs> ea;
    std::vector<LinkArgs> la;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        Plan& pl = plan[g];
        const int D = pl.D;
        const uint32_t* hc = ex.hsmall + g * SMALL_WORDS + SM_COUNTS;
        uint32_t* rk = v->ranks_.as<uint32_t>();
        LevelBuild LB[MAXLEV];
        for (int L = 0; L < D; L++) {
            LB[L].dom = pl.dom[L];
            LB[L].bmA = A(g, L);
            LB[L].bmB = pl.dilate ? B(g, L) : A(g, L);
            LB[L].rankA = rk + pl.rk_off[L];
            LB[L].rankE = pl.dilate ? rk + pl.rk_off[L] + pl.dom[L].nwords : nullptr;
            LB[L].countA = hc[2 * L];
            LB[L].countE = pl.dilate ? hc[2 * L + 1] : 0u;
            LB[L].lev = L;
            LB[L].range = int(v->range_[L]);
            u64 cnt = u64(LB[L].countA) + LB[L].countE;
            v->pool_[0][L].lastEle = v->pool_[0][L].usedNum = cnt;
            rc = v->ensure_pool_capacity(0, L, std::max<u64>(cnt, 1));
            if (rc) FAILG(g, rc);
            if (L > 0) {
                v->pool_[1][L].lastEle = v->pool_[1][L].usedNum = cnt;
                rc = v->ensure_pool_capacity(1, L, std::max<u64>(cnt, 1));
                if (rc) FAILG(g, rc);
            }
        }
        v->pool_[0][D].lastEle = v->pool_[0][D].usedNum = 1;
        v->pool_[1][D].lastEle = v->pool_[1][D].usedNum = 1;
        rc = v->ensure_pool_capacity(0, D, 1);
        if (rc) FAILG(g, rc);
        rc = v->ensure_pool_capacity(1, D, 1);
        if (rc) FAILG(g, rc);
        v->compute_top_lev();
        RootArgs r;
        r.root = v->pool_[0][D].dev.as<Node>();
        for (int a = 0; a < 3; a++) r.pos[a] = pl.rootPos[a];
        r.lev = D;
        ra.push_back(r);
        cf.push_back(FillArgs{v->pool_[1][D].dev.p, v->voxcnt(D), ID_UNDEF64, 1});
        // the brick table comes with the level-0 emission (one entry per box bit)
        u64* btab = nullptr;
        if (pl.dom[0].nbits <= BRICK_TAB_MAX) {
            XCHK(v, v->btab_.ensure(size_t(pl.dom[0].nbits) * 8));
            btab = v->btab_.as<u64>();
            v->btab_dom_ = pl.dom[0];
        }
        pl.btab = btab != nullptr;
        for (int L = 0; L < D; L++) {
            ea.push_back(EmitArgs{LB[L], v->pool_[0][L].dev.as<Node>(), L > 0 ? 1 : 0, L == 0 ? btab : nullptr});
            if (L > 0) cf.push_back(FillArgs{v->pool_[1][L].dev.p, v->pool_[1][L].lastEle * v->voxcnt(L), ID_UNDEF64, 1});
        }
        // gvdbLinkBricks walks from the top node; levels above it stay unlinked
        for (int L = 0; L < D; L++) {
            if (L + 1 > v->top_lev_) continue;
            const bool toRoot = (L + 1 == D);
            la.push_back(LinkArgs{LB[L], toRoot ? LB[L] : LB[L + 1], toRoot ? 1 : 0, v->logdim_[L + 1],
                                  v->pool_[0][L].dev.as<Node>(), v->pool_[1][L + 1].dev.as<u64>(), v->voxcnt(L + 1)});
        }
    }
    XCHK(v0, launch_write_root_multi(ra.data(), int(ra.size()), ex.ar, ex.s));
    XCHK(v0, launch_fill_multi(cf.data(), int(cf.size()), ex.ar, ex.s));
    XCHK(v0, launch_emit_nodes_multi(ea.data(), int(ea.size()), ex.ar, ex.s));
    XCHK(v0, launch_link_level_multi(la.data(), int(la.size()), ex.ar, ex.s));
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        v->root_ = Elem(0, plan[g].D, 0);
        v->rebuild_topo_ = false;
        v->host_valid_ = false;
        v->dev_stale_ = false;
        v->topo_gen_++;
        v->btab_gen_ = plan[g].btab ? v->topo_gen_ : ~0u;
    }
    return GVDB_OK;
}

// AccumulateTopology (gvdb_volume_gvdb.cpp:1158-1178): depth at least
// min_depth; RebuildTopology while a full rebuild is pending; otherwise the
// incremental build keeping the flags of the existing leaves.  A depth above
// the root's level is refused (the reference would walk past its root).
int Volume::accumulate_topology(int n, float radius, const float orig[3], int min_depth) {
    if (capture_) return not_recordable("AccumulateTopology");
    if (levels_ == 0) return fail(GVDB_ERR_STATE, "AccumulateTopology before Configure");
    if (n <= 0 || !pts_.base || n > pts_.n) return fail(GVDB_ERR_INVALID, "AccumulateTopology: point count exceeds SetPoints");
    if (!radius_ok(radius)) return fail(GVDB_ERR_UNSUPPORTED, "AccumulateTopology: radius must be in [0, 1024]");
    if (rebuild_topo_) return rebuild_topology(n, radius, orig);
    const Exec ex = self_exec();
    stage_begin(ST_TOPO);
    PointSrc P = pts_;
    P.n = n;
    XCHK(this, bbox_part_.ensure(size_t(BBOX_PART_WORDS) * 4));
    BboxArgs ba{P, {orig[0], orig[1], orig[2]}, ex.dsmall + SM_BBOX, bbox_part_.as<uint32_t>()};
    XCHK(this, launch_bbox_multi(&ba, 1, ex.ar, ex.s));
    int rc = drain(this, ex, 1);
    if (rc) return rc;
    float mnmx[6];
    for (int a = 0; a < 3; a++) {
        mnmx[a] = key_float(ex.hsmall[SM_BBOX + a]);
        mnmx[3 + a] = key_float(ex.hsmall[SM_BBOX + 3 + a]);
        posMin_[a] = mnmx[a] - 1;
        posMax_[a] = mnmx[3 + a] + 2;
        posMin_[a] = 0 > posMin_[a] ? 0 : posMin_[a];
        posMax_[a] = 0 > posMax_[a] ? 0 : posMax_[a];
    }
    int rootPos[3];
    int depth = determine_depth(rootPos);
    if (depth < min_depth) depth = min_depth;
    if (depth == 0) return fail(GVDB_ERR_DEPTH, "Points extents smaller than 1 brick. Not yet supported.");
    if (depth < 0 || depth > 5 || depth >= levels_) return fail(GVDB_ERR_DEPTH, "Point extents exceed 5 levels.");
    if (root_ == ID_UNDEFL || depth > int((root_ >> 8) & 0xFF))
        return fail(GVDB_ERR_DEPTH, "AccumulateTopology: the points need a deeper tree than the current one "
                                    "(call RequestFullRebuild)");
    rc = incremental_topology(n, radius, orig, depth, true, mnmx, ex);
    if (rc) return rc;
    stage_end(ST_TOPO);
    return GVDB_OK;
}

// ActivateIncreBricksGPU (gvdb_volume_gvdb.cpp:1020-1156) for one grid:
// markers, the bitmaps of missing nodes, and (only when some node is
// missing) markers applied, nodes appended in key order, levels linked
// top-down, unflagged leaves and childless nodes delinked bottom-up, the
// leaf used count updated.  `mnmx`: min and max of p + orig over the points.
int Volume::incremental_topology(int n, float radius, const float orig[3], int D, bool accum, const float* mnmx,
                                 const Exec& ex) {
    const u64 total = pool_[0][0].lastEle;
    if (total == 0) return GVDB_OK;
    int rc = commit_pools();
    if (rc) return rc;
    if (D <= 0 || D >= levels_) return fail(GVDB_ERR_DEPTH, "incremental topology: depth out of range");
    BitDomain dom[MAXLEV];
    u64 bm_off[MAXLEV], tmp_off[MAXLEV], bm_words = 0, tmp_words = 0;
    // boxes above 2^34 cells (or the keyed path forced): missing cells as
    // sorted keys instead of bitmaps
    bool keyed = false;
    for (int L = 0; L < D; L++) {
        const int rr = int(range_[L]);
        BitDomain& d = dom[L];
        d.nbits = 1;
        for (int a = 0; a < 3; a++) {
            const float lo = std::isfinite(mnmx[a]) ? std::max(mnmx[a], 0.0f) : 0.0f;
            const float hi = std::isfinite(mnmx[3 + a]) ? std::max(mnmx[3 + a], 0.0f) : 0.0f;
            const int blo = int(float(int(lo)) - radius), bhi = int(float(int(hi)) + radius);
            d.lo[a] = std::max(0, blo / rr);
            d.dim[a] = std::max(1, bhi / rr - d.lo[a] + 1);
            d.nbits *= u64(d.dim[a]);
        }
        if (keyed_build_wanted(d)) keyed = true;
        d.nwords = (d.nbits + 31) / 32;
        bm_off[L] = bm_words;
        bm_words += d.nwords;
        tmp_off[L] = tmp_words;
        tmp_words += scan_tmp_words(d.nwords);
    }
    if (!keyed) {
        XCHK(this, bitmaps_.ensure(bm_words * 4 + 256));
        XCHK(this, ranks_.ensure(bm_words * 4 + 256));
        XCHK(this, scan_tmp_.ensure(tmp_words * 4 + 256));
    }
    XCHK(this, incr_marker_.ensure(size_t(total) * 4));
    Tree t = make_tree();
    std::vector<FillArgs> fa;
    if (!keyed) fa.push_back(FillArgs{bitmaps_.p, bm_words, 0, 0});
    if (!accum) fa.push_back(FillArgs{incr_marker_.p, total, 0, 0});
    XCHK(this, launch_fill_multi(fa.data(), int(fa.size()), ex.ar, ex.s));
    if (accum) {
        FlagArgs fl{t, incr_marker_.as<uint32_t>(), int(total), 0};
        XCHK(this, launch_leaf_flags_multi(&fl, 1, ex.ar, ex.s));
    }
    IncreMarkArgs ma;
    memset(&ma, 0, sizeof(ma));
    ma.t = t;
    ma.pts = pts_;
    ma.pts.n = n;
    for (int a = 0; a < 3; a++) ma.o[a] = orig[a];
    ma.radius = radius;
    ma.D = D;
    for (int L = 0; L < D; L++) {
        ma.dom[L] = dom[L];
        ma.bm[L] = bitmaps_.as<uint32_t>() + bm_off[L];
    }
    ma.marker = incr_marker_.as<uint32_t>();
    u64 cnt[MAXLEV], first[MAXLEV], added = 0;
    if (!keyed) {
        XCHK(this, launch_incre_mark(ma, ex.s));
        std::vector<ScanArgs> sa;
        for (int L = 0; L < D; L++)
            sa.push_back(ScanArgs{bitmaps_.as<uint32_t>() + bm_off[L], nullptr, 1, ranks_.as<uint32_t>() + bm_off[L],
                                  dom[L].nwords, scan_tmp_.as<uint32_t>() + tmp_off[L], ex.dsmall + SM_COUNTS + 2 * L});
        XCHK(this, launch_scan_multi(sa.data(), int(sa.size()), ex.ar, ex.s));
        rc = drain(this, ex, 1);
        if (rc) return rc;
        for (int L = 0; L < D; L++) {
            cnt[L] = ex.hsmall[SM_COUNTS + 2 * L];
            first[L] = 0;
            added += cnt[L];
        }
    } else {
        // count (and markers), 64-bit total, offsets; then the keys, sorted
        // and made unique; each level's slice of the list
        XCHK(this, kb_cnt_.ensure(size_t(std::max(n, 1)) * 4));
        XCHK(this, kb_off_.ensure(size_t(std::max(n, 1)) * 4));
        XCHK(this, launch_incre_keys(ma, kb_cnt_.as<uint32_t>(), nullptr, nullptr, false, ex.s));
        XCHK(this, launch_sum_u32(kb_cnt_.as<uint32_t>(), u64(n), reinterpret_cast<u64*>(ex.dsmall + SM_LEN), ex.s));
        XCHK(this, scan_tmp_.ensure(scan_tmp_words(u64(n)) * 4 + 256));
        XCHK(this, launch_scan_u32(kb_cnt_.as<uint32_t>(), nullptr, 0, kb_off_.as<uint32_t>(), u64(n), nullptr,
                                   scan_tmp_.as<uint32_t>(), ex.s));
        rc = drain(this, ex, 1);
        if (rc) return rc;
        u64 nk = 0;
        memcpy(&nk, ex.hsmall + SM_LEN, 8);
        if (nk > (u64(1) << 31)) {
            kb_release();
            return fail(GVDB_ERR_RANGE, "incremental topology: more than 2^31 missing cells over the points");
        }
        XCHK(this, kb_keys_.ensure(size_t(std::max<u64>(nk, 1)) * 8));
        XCHK(this, launch_incre_keys(ma, kb_cnt_.as<uint32_t>(), kb_off_.as<uint32_t>(), kb_keys_.as<u64>(), true, ex.s));
        uint32_t nu = 0;
        rc = kb_sort_unique(kb_keys_.as<u64>(), nk, 59, ~0ull, kb_list_, &nu, ex);
        if (rc) return rc;
        FillArgs z{ex.dsmall + SM_COUNTS, 2 * MAXLEV, 0, 0};
        XCHK(this, launch_fill_multi(&z, 1, ex.ar, ex.s));
        XCHK(this, launch_level_counts(kb_list_.as<u64>(), nu, ex.dsmall + SM_COUNTS, ex.s));
        rc = drain(this, ex, 1);
        if (rc) return rc;
        for (int L = 0; L < D; L++) {
            first[L] = ex.hsmall[SM_COUNTS + 2 * L];
            cnt[L] = ex.hsmall[SM_COUNTS + 2 * L + 1] - first[L];
            added += cnt[L];
        }
        if (added == 0) kb_release();
    }
    if (added == 0) return GVDB_OK;   // the reference returns before applying the markers
    XCHK(this, launch_apply_markers(pool_[0][0].dev.as<Node>(), incr_marker_.as<uint32_t>(), total, ex.s));
    std::vector<FillArgs> cf;
    u64 base0[MAXLEV], base1[MAXLEV];
    for (int L = 0; L < D; L++) {
        base0[L] = pool_[0][L].lastEle;
        base1[L] = L > 0 ? pool_[1][L].lastEle : 0;
        rc = ensure_pool_capacity(0, L, std::max<u64>(base0[L] + cnt[L], 1), true, ex.s);
        if (rc) return rc;
        if (L > 0) {
            rc = ensure_pool_capacity(1, L, std::max<u64>(base1[L] + cnt[L], 1), true, ex.s);
            if (rc) return rc;
        }
    }
    for (int L = 0; L < D; L++) {
        if (cnt[L] == 0) continue;
        if (keyed) {
            XCHK(this, launch_incre_emit_list(kb_list_.as<u64>() + first[L], uint32_t(cnt[L]), dom[L],
                                              pool_[0][L].dev.as<Node>(), L, int(range_[L]), L > 0 ? 1 : 0, base0[L],
                                              base1[L], ex.s));
        } else {
            IncreEmitArgs ea{dom[L], bitmaps_.as<uint32_t>() + bm_off[L], ranks_.as<uint32_t>() + bm_off[L],
                             pool_[0][L].dev.as<Node>(), L, int(range_[L]), L > 0 ? 1 : 0, base0[L], base1[L]};
            XCHK(this, launch_incre_emit(ea, ex.s));
        }
        if (L > 0)
            cf.push_back(FillArgs{pool_[1][L].dev.as<u64>() + base1[L] * voxcnt(L), cnt[L] * voxcnt(L), ID_UNDEF64, 1});
        pool_[0][L].lastEle += cnt[L];
        pool_[0][L].usedNum += cnt[L];
        if (L > 0) {
            pool_[1][L].lastEle += cnt[L];
            pool_[1][L].usedNum += cnt[L];
        }
    }
    XCHK(this, launch_fill_multi(cf.data(), int(cf.size()), ex.ar, ex.s));
    compute_top_lev();
    t = make_tree();
    for (int L = D - 1; L >= 0; L--) XCHK(this, launch_link_walk(t, L, ex.s));
    XCHK(this, launch_delink_leaves(t, ex.s));
    for (int L = 1; L < D; L++) XCHK(this, launch_delink_level(t, L, ex.s));
    FillArgs z{ex.dsmall + SM_UNFLAGGED, 1, 0, 0};
    XCHK(this, launch_fill_multi(&z, 1, ex.ar, ex.s));
    XCHK(this, launch_count_unflagged(t, ex.dsmall + SM_UNFLAGGED, ex.s));
    rc = drain(this, ex, 1);
    if (rc) return rc;
    if (keyed) kb_release();
    pool_[0][0].usedNum = pool_[0][0].lastEle - ex.hsmall[SM_UNFLAGGED];
    compute_top_lev();
    host_valid_ = false;
    dev_stale_ = false;
    topo_gen_++;
    return GVDB_OK;
}

// -------------------------------------------------------- FinishTopology ----
int Volume::multi_finish_topology(Volume* const* vs, int G, bool commit, bool bounds, const Exec& ex, int* failed) {
    if (ex.cap) FAILG(0, vs[0]->not_recordable("FinishTopology"));
    for (int g = 0; g < G; g++)
        if (vs[g]->levels_ == 0) FAILG(g, vs[g]->fail(GVDB_ERR_STATE, "FinishTopology before Configure"));
    if (bounds) {
        // ComputeBounds (gvdb_volume_gvdb.cpp:1792-1816)
        std::vector<BoundsArgs> bd;
        std::vector<int> who;
        for (int g = 0; g < G; g++) {
            Volume* v = vs[g];
            Pool& p = v->pool_[0][0];
            float mnv[3] = {0, 0, 0}, mxv[3] = {0, 0, 0};
            if (p.lastEle && v->dev_stale_) {
                // host-built pools: bounds from the host mirror
                Node* c = v->hnode(Elem(0, 0, 0));
                for (int a = 0; a < 3; a++) mnv[a] = mxv[a] = float(c->mPos[a]);
                for (u64 i = 0; i < p.lastEle; i++) {
                    c = v->hnode(Elem(0, 0, i));
                    if (!c->mFlags) continue;
                    for (int a = 0; a < 3; a++) {
                        if (c->mPos[a] < mnv[a]) mnv[a] = float(c->mPos[a]);
                        if (c->mPos[a] + v->range_[0] > mxv[a]) mxv[a] = float(c->mPos[a] + v->range_[0]);
                    }
                }
            } else if (p.lastEle) {
                bd.push_back(BoundsArgs{v->make_tree(), int(p.lastEle),
                                        reinterpret_cast<int*>(ex.dsmall + g * SMALL_WORDS + SM_BOUNDS)});
                who.push_back(g);
            }
            for (int a = 0; a < 3; a++) { v->objMin_[a] = mnv[a]; v->objMax_[a] = mxv[a]; }
            v->bounds_pending_ = false;
        }
        if (!bd.empty()) {
            // no host wait: each grid's bounds come back behind an event and
            // are taken when read (the render, the user-kernel view, GetInfo)
            XCHK(vs[0], launch_leaf_bounds_multi(bd.data(), int(bd.size()), ex.ar, ex.s));
            for (int g : who) {
                Volume* v = vs[g];
                XCHK(v, hipMemcpyAsync(v->pinned_ + PIN_BOUNDS, ex.dsmall + g * SMALL_WORDS + SM_BOUNDS, 6 * 4,
                                       hipMemcpyDeviceToHost, ex.s));
                XCHK(v, hipEventRecord(v->bounds_ev_, ex.s));
                v->bounds_pending_ = true;
            }
        }
    }
    for (int g = 0; g < G; g++) {
        if (commit) {
            int rc = vs[g]->commit_pools();
            if (rc) FAILG(g, rc);
        }
        vs[g]->compute_top_lev();
    }
    return GVDB_OK;
}

// ----------------------------------------------------------- UpdateAtlas ----
int Volume::multi_update_atlas(Volume* const* vs, int G, const Exec& ex, int* failed) {
    if (ex.cap) FAILG(0, vs[0]->not_recordable("UpdateAtlas"));
    std::vector<FillArgs> fa;
    std::vector<FlagArgs> fl;
    std::vector<ScanArgs> sa;
    std::vector<AtlasArgs> aa;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (!v->chan_[0].exists) FAILG(g, v->fail(GVDB_ERR_STATE, "UpdateAtlas: no atlas created"));
        int rc = v->commit_pools();
        if (rc) FAILG(g, rc);
        if (G == 1) v->stage_begin(ST_ATLAS);
        // UpdateAtlas (gvdb_volume_gvdb.cpp:2630-2704)
        u64 total = v->pool_[0][0].lastEle, used = v->pool_[0][0].usedNum;
        v->atlas_used_ = 0;
        if (used > v->atlas_max_) {
            int cz = int(ceil(used / float(v->atlas_cnt_[0] * v->atlas_cnt_[1])));
            rc = v->resize_atlas(cz);
            if (rc) FAILG(g, rc);
        }
        v->amap_max_ = u64(v->atlas_cnt_[0]) * v->atlas_cnt_[1] * v->atlas_cnt_[2];
        XCHK(v, v->amap_.ensure(size_t(v->amap_max_) * sizeof(AtlasNode)));
        fa.push_back(FillArgs{v->amap_.p, v->amap_max_ * 4, 0xFFFFFFFFull, 0});
        if (total) {
            XCHK(v, v->flags_.ensure(size_t(total) * 4));
            XCHK(v, v->slot_scan_.ensure(size_t(total) * 4));
            XCHK(v, v->scan_tmp_.ensure(scan_tmp_words(total) * 4 + 256));
            Tree t = v->make_tree();
            fl.push_back(FlagArgs{t, v->flags_.as<uint32_t>(), int(total), 0});
            sa.push_back(ScanArgs{v->flags_.as<uint32_t>(), nullptr, 0, v->slot_scan_.as<uint32_t>(), total,
                                  v->scan_tmp_.as<uint32_t>(), nullptr});
            aa.push_back(AtlasArgs{t, v->slot_scan_.as<uint32_t>(), int(total)});
        }
    }
    Volume* v0 = vs[0];
    XCHK(v0, launch_fill_multi(fa.data(), int(fa.size()), ex.ar, ex.s));
    XCHK(v0, launch_leaf_flags_multi(fl.data(), int(fl.size()), ex.ar, ex.s));
    XCHK(v0, launch_scan_multi(sa.data(), int(sa.size()), ex.ar, ex.s));
    XCHK(v0, launch_update_atlas_multi(aa.data(), int(aa.size()), ex.ar, ex.s));
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        v->atlas_used_ = v->pool_[0][0].usedNum;     // the flagged leaves (AtlasAlloc calls)
        v->host_valid_ = false;
        v->compute_top_lev();
        v->atlas_gen_ = v->topo_gen_;
        v->atlas_epoch_++;
    }
    if (G == 1) v0->stage_end(ST_ATLAS);
    return GVDB_OK;
}

int Volume::multi_clear_channel(Volume* const* vs, int G, int c, const Exec& ex, int* failed) {
    // ClearChannel -> AtlasFill (gvdb_allocator.cpp:813-822)
    std::vector<FillArgs> fa;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (c < 0 || c >= 8 || !v->chan_[c].exists) FAILG(g, v->fail(GVDB_ERR_INVALID, "ClearChannel: no such channel"));
        fa.push_back(FillArgs{v->chan_[c].data.p, v->chan_[c].data.bytes / 4, 0, 0});
        v->atlas_epoch_++;
        if (ex.cap) ex.cap->note(v);
    }
    if (G == 1) vs[0]->stage_begin(ST_CLEAR);   // (no events while recording)
    XCHK(vs[0], launch_fill_multi(fa.data(), G, ex.ar, ex.s));
    if (G == 1) vs[0]->stage_end(ST_CLEAR);
    return GVDB_OK;
}

// --------------------------------------------------- InsertPointsSubcell ----
int Volume::multi_insert(Volume* const* vs, int G, int scs, const int* n, const float* radius, const float* trans3,
                         int* len, const Exec& ex, int* failed, int rule) {
    // InsertPointsSubcell (gvdb_volume_gvdb.cpp:4971-5092); with rule BIN_FP16
    // the binning of InsertPointsSubcell_FP16 (:5094-5189)
    struct Plan {
        int sd, spb, R;
        u64 numSC, numMap;
        bool pull;
        uint32_t linked = 0, total = 0, orphans = 0;
        PointSrc P;
        Tree t;
    };
    std::vector<Plan> plan(G);
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        Plan& pl = plan[g];
        v->sc16_valid_ = false;
        u64 numBrick = v->pool_[0][0].usedNum;
        if (numBrick == 0) FAILG(g, v->fail(GVDB_ERR_NOBRICKS, "Warning: InsertPointsSubcell yet no bricks exist."));
        if (scs <= 0 || v->res_[0] % scs != 0)
            FAILG(g, v->fail(GVDB_ERR_INVALID, "InsertPointsSubcell: subcell size must divide the brick"));
        if (n[g] < 0 || n[g] > v->pts_.n || (n[g] > 0 && !v->pts_.base))
            FAILG(g, v->fail(GVDB_ERR_INVALID, "InsertPointsSubcell: point count exceeds SetPoints"));
        if (v->clr_.base && n[g] > v->clr_.n)
            FAILG(g, v->fail(GVDB_ERR_INVALID, "InsertPointsSubcell: point count exceeds SetPoints' colours"));
        v->bkt_clr_ = false;
        if (!radius_ok(radius[g]))
            FAILG(g, v->fail(GVDB_ERR_UNSUPPORTED, "InsertPointsSubcell: radius must be in [0, 1024]"));
        if (v->atlas_gen_ != v->topo_gen_)
            FAILG(g, v->fail(GVDB_ERR_STATE, "InsertPointsSubcell: call UpdateAtlas after the topology changes"));
        int rc = v->commit_pools();
        if (rc) FAILG(g, rc);
        pl.sd = v->res_[0] / scs;
        pl.spb = pl.sd * pl.sd * pl.sd;
        pl.numSC = u64(pl.spb) * numBrick;
        pl.numMap = v->pool_[0][0].lastEle;
        if (pl.numSC > 0xFFFFFFFFull) FAILG(g, v->fail(GVDB_ERR_RANGE, "InsertPointsSubcell: too many subcells"));
        // pull path when the neighbourhood of a brick stays inside its 27 neighbours
        pl.R = int(ceilf((radius[g] + 1.0f) / float(scs)));
        pl.pull = pl.R <= pl.sd && pull_cells(pl.sd, pl.R) <= pull_max_cells() && pl.spb <= pull_max_spb();
        if (ex.cap) {
            if (!pl.pull || rule != BIN_FLOAT)
                FAILG(g, v->not_recordable("InsertPointsSubcell outside the bucketed path"));
            ex.cap->note(v);
        }
        XCHK(v, v->sc_flags_.ensure(size_t(pl.numMap) * 4));
        XCHK(v, v->sc_map_.ensure(size_t(pl.numMap) * 4));
        XCHK(v, v->sc_cnt_.ensure(size_t(pl.numSC) * 4));
        XCHK(v, v->sc_use_.ensure(size_t(pl.numSC) * 4));
        XCHK(v, v->sc_off_.ensure(size_t(pl.numSC) * 4));
        u64 nb = pl.numMap * u64(pl.spb);
        // scan scratch: mapping, then buckets (pull) or counts, in separate regions
        XCHK(v, v->scan_tmp_.ensure((scan_tmp_words(pl.numMap) + scan_tmp_words(std::max(nb, pl.numSC))) * 4 + 512));
        pl.P = v->pts_;
        pl.P.n = n[g];
        if (pl.pull) {
            XCHK(v, v->nbr_.ensure(size_t(pl.numMap) * 27 * 4));
            XCHK(v, v->bkt_cnt_.ensure(size_t(nb) * 4));
            XCHK(v, v->bkt_off_.ensure(size_t(nb) * 4));
            XCHK(v, v->bkt_key_.ensure(size_t(std::max(n[g], 1)) * 4));
            XCHK(v, v->bkt_slot_.ensure(size_t(std::max(n[g], 1)) * 4));
            XCHK(v, v->sorted_.ensure(size_t(std::max(n[g], 1)) * 16));
        }
        pl.t = v->make_tree();
    }
    Volume* v0 = vs[0];
    if (G == 1) v0->stage_begin(ST_INSERT);
    // one grid, not recording: the small words come back behind an event
    // recorded before the last kernels, not after a wait for the stream
    const bool early = G == 1 && !ex.cap;
    auto small = [&](int g, int k) { return ex.dsmall + g * SMALL_WORDS + k; };
    auto tmp2 = [&](int g) { return vs[g]->scan_tmp_.as<uint32_t>() + scan_tmp_words(plan[g].numMap) + 64; };
    {
        std::vector<FlagArgs> fl;
        std::vector<ScanArgs> sa;
        std::vector<NbrArgs> na;
        std::vector<FillArgs> fa;
        for (int g = 0; g < G; g++) {
            Volume* v = vs[g];
            Plan& pl = plan[g];
            fl.push_back(FlagArgs{pl.t, v->sc_flags_.as<uint32_t>(), int(pl.numMap), 1});
            sa.push_back(ScanArgs{v->sc_flags_.as<uint32_t>(), nullptr, 0, v->sc_map_.as<uint32_t>(), pl.numMap,
                                  v->scan_tmp_.as<uint32_t>(), small(g, SM_LINKED)});
            fa.push_back(FillArgs{v->sc_cnt_.p, pl.numSC, 0, 0});
            if (pl.pull) {
                if (v->nbr_gen_ != v->topo_gen_) {
                    // k_neighbors writes every brick's mask whole
                    XCHK(v, v->nmask_.ensure(size_t(pl.numMap) * 4));
                    XCHK(v, v->nbase_.ensure(size_t(pl.numMap) * 27 * 8));
                    na.push_back(NbrArgs{pl.t, v->nbr_.as<int>(), int(pl.numMap), v->nmask_.as<uint32_t>(),
                                         v->nbase_.as<long long>(), v->brick_tab()});
                    v->nbr_gen_ = v->topo_gen_;
                }
                fa.push_back(FillArgs{v->bkt_cnt_.p, pl.numMap * u64(pl.spb), 0, 0});
                fa.push_back(FillArgs{small(g, SM_ORPHANS), 1, 0, 0});
                XCHK(v, v->lenparts_.ensure(64 * 8));
                fa.push_back(FillArgs{v->lenparts_.p, 64, 0, 1});
            }
        }
        XCHK(v0, launch_leaf_flags_multi(fl.data(), int(fl.size()), ex.ar, ex.s));
        XCHK(v0, launch_scan_multi(sa.data(), int(sa.size()), ex.ar, ex.s));
        XCHK(v0, launch_neighbors_multi(na.data(), int(na.size()), ex.ar, ex.s));
        XCHK(v0, launch_fill_multi(fa.data(), int(fa.size()), ex.ar, ex.s));
    }
    {
        std::vector<BucketArgs> ba;
        std::vector<ScanArgs> sb, sc;
        std::vector<ScatterArgs> sca;
        std::vector<const u64*> lp;
        std::vector<u64*> lo;
        for (int g = 0; g < G; g++) {
            Volume* v = vs[g];
            Plan& pl = plan[g];
            const float* tr = trans3 + 3 * g;
            if (!pl.pull) {
                // push path: the reference's count pass with global atomics
                XCHK(v, launch_count_subcell(pl.t, pl.P, scs, pl.spb, tr, radius[g], v->sc_map_.as<uint32_t>(),
                                             v->sc_cnt_.as<uint32_t>(), rule, ex.s));
                XCHK(v, launch_sum_u32(v->sc_cnt_.as<uint32_t>(), pl.numSC, reinterpret_cast<u64*>(small(g, SM_LEN)),
                                       ex.s));
                sc.push_back(ScanArgs{v->sc_cnt_.as<uint32_t>(), nullptr, 0, v->sc_off_.as<uint32_t>(), pl.numSC,
                                      tmp2(g), small(g, SM_TOTAL)});
                continue;
            }
            u64 nb = pl.numMap * u64(pl.spb);
            ba.push_back(BucketArgs{pl.t, pl.P, scs, pl.spb, {tr[0], tr[1], tr[2]}, v->bkt_key_.as<uint32_t>(),
                                    v->bkt_slot_.as<uint32_t>(), v->bkt_cnt_.as<uint32_t>(), small(g, SM_ORPHANS),
                                    v->nmask_.as<uint32_t>(), radius[g], v->lenparts_.as<u64>(), rule, v->brick_tab()});
            lp.push_back(v->lenparts_.as<u64>());
            lo.push_back(reinterpret_cast<u64*>(small(g, SM_LEN)));
            sb.push_back(ScanArgs{v->bkt_cnt_.as<uint32_t>(), nullptr, 0, v->bkt_off_.as<uint32_t>(), nb, tmp2(g),
                                  small(g, SM_BUCKETS)});
            // SetPoints' colours travel with the bucketed points (fourth word)
            const bool clr = v->clr_.base && rule == BIN_FLOAT;
            sca.push_back(ScatterArgs{pl.P, clr ? v->clr_ : PointSrc{nullptr, 0, 0, 4}, {tr[0], tr[1], tr[2]},
                                      v->bkt_key_.as<uint32_t>(), v->bkt_slot_.as<uint32_t>(),
                                      v->bkt_off_.as<uint32_t>(), v->sorted_.as<float4>()});
            v->bkt_clr_ = clr;
        }
        XCHK(v0, launch_bucket_points_multi(ba.data(), int(ba.size()), ex.ar, ex.s));
        XCHK(v0, launch_sum_parts_multi(lp.data(), lo.data(), int(lp.size()), ex.ar, ex.s));
        if (early) {
            // every word the host reads below is final here: it comes back
            // while the bucket scan and scatter still run, and the gather the
            // caller launches next queues behind them
            XCHK(v0, hipMemcpyAsync(ex.hsmall, ex.dsmall, SMALL_WORDS * 4, hipMemcpyDeviceToHost, ex.s));
            XCHK(v0, hipEventRecord(v0->small_ev_, ex.s));
        }
        XCHK(v0, launch_scan_multi(sb.data(), int(sb.size()), ex.ar, ex.s));
        XCHK(v0, launch_bucket_scatter_multi(sca.data(), int(sca.size()), ex.ar, ex.s));
        XCHK(v0, launch_scan_multi(sc.data(), int(sc.size()), ex.ar, ex.s));
    }
    int rc = GVDB_OK;
    if (early)
        XCHK(v0, hipEventSynchronize(v0->small_ev_));
    else
        rc = drain(v0, ex, G);
    if (rc) return rc;
    if (ex.cap) {
        // recording: the entry counts arrive with the replay (graph_synchronize
        // reads them from the recorded copy); the calls between see the
        // previous count, which a bucketed gather does not read
        for (int g = 0; g < G; g++) {
            Volume* v = vs[g];
            Plan& pl = plan[g];
            len[g] = v->sc_len_;
            v->num_sc_ = int(pl.numSC);
            v->sc_map_n_ = int(pl.numMap);
            v->sc_gen_ = v->topo_gen_;
            v->sc16_valid_ = false;
            v->last_insert_pull_ = true;
            v->entries_stale_ = true;
            v->bkt_scs_ = scs;
            v->bkt_R_ = pl.R;
            v->bkt_radius_ = radius[g];
            v->bkt_gen_ = v->topo_gen_;
            ex.cap->inserts.push_back(Graph::Insert{v, ex.hsmall + size_t(g) * SMALL_WORDS});
        }
        return GVDB_OK;
    }
    // grids with points outside every brick fall back to the push path
    bool redo = false;
    for (int g = 0; g < G; g++) {
        Plan& pl = plan[g];
        pl.linked = ex.hsmall[g * SMALL_WORDS + SM_LINKED];
        pl.orphans = pl.pull ? ex.hsmall[g * SMALL_WORDS + SM_ORPHANS] : 0;
        {
            // the exact entry total in 64 bits (the u32 scan's wraps)
            u64 l = 0;
            memcpy(&l, ex.hsmall + g * SMALL_WORDS + SM_LEN, 8);
            pl.total = l > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(l);
        }
        if (pl.pull && pl.orphans) {
            Volume* v = vs[g];
            pl.pull = false;
            redo = true;
            XCHK(v, launch_fill_u32(v->sc_cnt_.as<uint32_t>(), 0u, pl.numSC, ex.s));
            XCHK(v, launch_count_subcell(pl.t, pl.P, scs, pl.spb, trans3 + 3 * g, radius[g], v->sc_map_.as<uint32_t>(),
                                         v->sc_cnt_.as<uint32_t>(), rule, ex.s));
            XCHK(v, launch_sum_u32(v->sc_cnt_.as<uint32_t>(), pl.numSC, reinterpret_cast<u64*>(small(g, SM_LEN)),
                                   ex.s));
            XCHK(v, launch_scan_u32(v->sc_cnt_.as<uint32_t>(), nullptr, 0, v->sc_off_.as<uint32_t>(), pl.numSC,
                                    small(g, SM_TOTAL), tmp2(g), ex.s));
        }
    }
    if (redo) {
        rc = drain(v0, ex, G);
        if (rc) return rc;
        for (int g = 0; g < G; g++)
            if (!plan[g].pull) {
                u64 l = 0;
                memcpy(&l, ex.hsmall + g * SMALL_WORDS + SM_LEN, 8);
                plan[g].total = l > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(l);
            }
    }
    std::vector<SubPosArgs> sp;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        Plan& pl = plan[g];
        if (pl.total > 0x7FFFFFFFu) FAILG(g, v->fail(GVDB_ERR_RANGE, "InsertPointsSubcell: more than 2^31 subcell entries"));
        len[g] = int(pl.total);
        v->sc_len_ = int(pl.total);
        v->num_sc_ = int(pl.numSC);
        v->sc_bricks_ = int(pl.linked);
        v->sc_map_n_ = int(pl.numMap);
        v->sc_gen_ = v->topo_gen_;
        v->sc16_valid_ = false;
        v->last_insert_pull_ = pl.pull;
        // the pull path leaves the subcell tables to be built on demand
        // (materialize_entries); the gather reads the buckets directly
        v->entries_stale_ = pl.pull;
        v->bkt_scs_ = scs;
        v->bkt_R_ = pl.R;
        v->bkt_radius_ = radius[g];
        v->bkt_gen_ = pl.pull ? v->topo_gen_ : ~0u;
        if (pl.pull) continue;
        // an empty list still gets its subcell corners: GatherLevelSet writes
        // every voxel of the linked bricks
        XCHK(v, v->sc_pos_.ensure(size_t(pl.numSC) * 12));
        XCHK(v, v->sc_nid_.ensure(size_t(pl.numSC) * 4));
        if (pl.total > 0) {
            XCHK(v, v->sc_pnt_.ensure(size_t(pl.total) * 12));
            XCHK(v, launch_fill_u32(v->sc_cnt_.as<uint32_t>(), 0u, pl.numSC, ex.s));
            XCHK(v, launch_insert_subcell(pl.t, pl.P, scs, pl.spb, trans3 + 3 * g, radius[g], v->sc_map_.as<uint32_t>(),
                                          v->sc_cnt_.as<uint32_t>(), v->sc_off_.as<uint32_t>(), v->sc_pnt_.as<float>(),
                                          rule, ex.s));
        }
        // unordered entries: the gather considers all of them
        XCHK(v, hipMemcpyAsync(v->sc_use_.p, v->sc_cnt_.p, size_t(pl.numSC) * 4, hipMemcpyDeviceToDevice, ex.s));
        sp.push_back(SubPosArgs{pl.t, scs, pl.spb, v->sc_map_.as<uint32_t>(), v->sc_pos_.as<int>(), v->sc_nid_.as<int>(),
                                int(pl.numMap)});
    }
    XCHK(v0, launch_subcell_pos_multi(sp.data(), int(sp.size()), ex.ar, ex.s));
    if (G == 1) v0->stage_end(ST_INSERT);
    return GVDB_OK;
}

// Subcell tables of pull-path inserts, built when something reads them: the
// per-brick count pass, the offsets, the entries (useful ones first) and the
// subcell corners.
int Volume::multi_materialize(Volume* const* vs, int G, const Exec& ex, int* failed) {
    if (ex.cap) FAILG(0, vs[0]->not_recordable("building the subcell entry tables"));
    std::vector<FillArgs> fa;
    std::vector<PullArgs> pc, pw;
    std::vector<ScanArgs> sc;
    std::vector<SubPosArgs> sp;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (!v->entries_stale_) continue;
        if (v->bkt_gen_ != v->topo_gen_) FAILG(g, v->fail(GVDB_ERR_STATE, "subcell tables: topology changed since InsertPointsSubcell"));
        const int scs = v->bkt_scs_;
        const int sd = v->res_[0] / scs, spb = sd * sd * sd;
        const u64 numSC = u64(v->num_sc_), numMap = u64(v->sc_map_n_);
        XCHK(v, v->sc_pnt_.ensure(size_t(v->sc_len_) * 12));
        XCHK(v, v->sc_pos_.ensure(size_t(numSC) * 12));
        XCHK(v, v->sc_nid_.ensure(size_t(numSC) * 4));
        XCHK(v, v->scan_tmp_.ensure(scan_tmp_words(numSC) * 4 + 256));
        Tree t = v->make_tree();
        fa.push_back(FillArgs{v->sc_cnt_.p, numSC, 0, 0});
        PullArgs a{t, int(numMap), scs, v->bkt_R_, v->bkt_radius_, v->nbr_.as<int>(), v->bkt_cnt_.as<uint32_t>(),
                   v->bkt_off_.as<uint32_t>(), v->sorted_.as<float4>(), v->sc_map_.as<uint32_t>(),
                   v->sc_cnt_.as<uint32_t>(), v->sc_use_.as<uint32_t>(), v->sc_off_.as<uint32_t>(), nullptr};
        pc.push_back(a);
        a.sc_pnt = v->sc_pnt_.as<float>();
        pw.push_back(a);
        sc.push_back(ScanArgs{v->sc_cnt_.as<uint32_t>(), nullptr, 0, v->sc_off_.as<uint32_t>(), numSC,
                              v->scan_tmp_.as<uint32_t>(), nullptr});
        sp.push_back(SubPosArgs{t, scs, spb, v->sc_map_.as<uint32_t>(), v->sc_pos_.as<int>(), v->sc_nid_.as<int>(),
                                int(numMap)});
        v->entries_stale_ = false;
    }
    if (pc.empty()) return GVDB_OK;
    Volume* v0 = vs[0];
    XCHK(v0, launch_fill_multi(fa.data(), int(fa.size()), ex.ar, ex.s));
    XCHK(v0, launch_pull_multi(0, pc.data(), int(pc.size()), ex.ar, ex.s));
    XCHK(v0, launch_scan_multi(sc.data(), int(sc.size()), ex.ar, ex.s));
    XCHK(v0, launch_pull_multi(1, pw.data(), int(pw.size()), ex.ar, ex.s));
    XCHK(v0, launch_subcell_pos_multi(sp.data(), int(sp.size()), ex.ar, ex.s));
    return GVDB_OK;
}

// --------------------------------------------------------- GatherDensity ----
int Volume::multi_gather(Volume* const* vs, int G, int scs, const int* len, const float* radius, int chanD, bool accum,
                         bool levelset, const Exec& ex, int* failed, bool fp16, int chanC) {
    // GatherDensity (gvdb_volume_gvdb.cpp:5318-5344) or GatherLevelSet (:5348-5371,
    // which does not return early on an empty entry list)
    std::vector<GatherArgs> ga;
    std::vector<BucketGatherArgs> bg;
    std::vector<ColorArgs> ca;
    std::vector<Volume*> need;
    std::vector<int> need_g, ents;   // grid index of each volume to materialize / to gather from entries
    const char* what = fp16 ? "GatherLevelSet_FP16" : (levelset ? "GatherLevelSet" : "GatherDensity");
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if ((len[g] == 0 && !levelset) || v->pool_[0][0].usedNum == 0) continue;
        if (chanD < 0 || chanD >= 8 || !v->chan_[chanD].exists || v->chan_[chanD].type == T_UCHAR4)
            FAILG(g, v->fail(GVDB_ERR_INVALID, std::string(what) + ": no such float or half channel"));
        if (len[g] != v->sc_len_ || scs <= 0 || v->res_[0] % scs)
            FAILG(g, v->fail(GVDB_ERR_STATE, std::string(what) + ": call InsertPointsSubcell first"));
        if (v->atlas_gen_ != v->topo_gen_ || v->sc_gen_ != v->topo_gen_)
            FAILG(g, v->fail(GVDB_ERR_STATE, std::string(what) + ": topology changed since UpdateAtlas/InsertPointsSubcell"));
        int rc = v->commit_pools();
        if (rc) FAILG(g, rc);
        if (fp16 && !v->sc16_valid_)
            FAILG(g, v->fail(GVDB_ERR_STATE, std::string(what) + ": call InsertPointsSubcell_FP16 first"));
        const bool buckets = !fp16 && v->bkt_gen_ == v->topo_gen_ && scs == v->bkt_scs_ && radius[g] == v->bkt_radius_ &&
                             v->res_[0] == 8 && scs == 4 && v->bkt_R_ >= 1 && v->bkt_R_ <= 2;
        if (ex.cap) {
            if (!buckets) FAILG(g, v->not_recordable(what));
            ex.cap->note(v);
        }
        // the colour channel (gvdbGatherDensity's chanClr): only when the
        // insert carried SetPoints' colours, as the reference writes it only
        // when its per-entry colour table exists
        const bool colour = chanC >= 0 && v->bkt_clr_ && !fp16;
        if (colour) {
            if (chanC >= 8 || !v->chan_[chanC].exists || v->chan_[chanC].type != T_UCHAR4 || chanC == chanD)
                FAILG(g, v->fail(GVDB_ERR_INVALID, std::string(what) + ": the colour channel must be a uchar4 channel"));
            if (!buckets)
                FAILG(g, v->fail(GVDB_ERR_UNSUPPORTED, std::string(what) + ": colours need the bucketed insert at the "
                                                       "gather's radius (8^3 bricks, subcell 4)"));
            ca.push_back(ColorArgs{v->make_tree(), int(v->pool_[0][0].lastEle), v->bkt_R_, radius[g], scs,
                                   v->nbr_.as<int>(), v->bkt_cnt_.as<uint32_t>(), v->bkt_off_.as<uint32_t>(),
                                   v->sorted_.as<float4>(), v->chan_[chanC].data.p, accum ? 1 : 0});
        }
        v->atlas_epoch_++;
        if (buckets) {
            // the scatter kernel (no accumulate) also leaves the bricks' faces for the apron
            void* faces = nullptr;
            if (!accum && !levelset) {
                const size_t need = size_t(v->pool_[0][0].lastEle) * FACE_VALS * v->chan_[chanD].elem();
                if (v->faces_.ensure(need) == hipSuccess) {
                    faces = v->faces_.p;
                    v->faces_epoch_ = v->atlas_epoch_;
                    v->faces_topo_ = v->topo_gen_;
                    v->faces_chan_ = chanD;
                }
            }
            bg.push_back(BucketGatherArgs{v->make_tree(), int(v->pool_[0][0].lastEle), v->bkt_R_, radius[g],
                                          v->nbr_.as<int>(), v->bkt_cnt_.as<uint32_t>(), v->bkt_off_.as<uint32_t>(),
                                          v->sorted_.as<float4>(), v->chan_[chanD].data.p, v->chan_[chanD].type,
                                          accum ? 1 : 0, faces, levelset ? 1 : 0});
            continue;
        }
        if (v->entries_stale_) {
            need.push_back(v);
            need_g.push_back(g);
        }
        ents.push_back(g);
    }
    if (ents.empty() && bg.empty() && ca.empty()) return GVDB_OK;
    Volume* v0 = vs[0];
    if (!need.empty()) {
        // allocates the entry tables, so the entry pointers are taken only after it
        int fi = -1;
        int rc = multi_materialize(need.data(), int(need.size()), ex, &fi);
        if (rc) {
            if (failed && fi >= 0) *failed = need_g[fi];
            return rc;
        }
    }
    for (int g : ents) {
        Volume* v = vs[g];
        if (v->sc_len_ > 0 && !(fp16 ? v->sc_pnt16_.p : v->sc_pnt_.p))
            FAILG(g, v->fail(GVDB_ERR_STATE, std::string(what) + ": no subcell entries"));
        // a bucketed insert places first the entries that can raise a voxel
        // at its own radius; a gather at a larger radius reads them all
        const bool all = v->last_insert_pull_ && !(radius[g] <= v->bkt_radius_);
        ga.push_back(GatherArgs{v->make_tree(), scs, radius[g], v->sc_nid_.as<int>(), v->sc_cnt_.as<uint32_t>(),
                                all ? v->sc_cnt_.as<uint32_t>() : v->sc_use_.as<uint32_t>(), v->sc_off_.as<uint32_t>(),
                                v->sc_pos_.as<int>(),
                                v->sc_pnt_.as<float>(), v->chan_[chanD].data.p, v->chan_[chanD].type, accum ? 1 : 0,
                                v->sc_bricks_, levelset ? 1 : 0});
        if (fp16) {
            ga.back().sc_pnt16 = v->sc_pnt16_.as<uint16_t>();
            for (int a = 0; a < 3; a++) {
                ga.back().qmin[a] = v->q_min_[a];
                ga.back().qrange[a] = v->q_range_[a];
            }
        }
    }
    const bool timed = G == 1 && !ex.cap;   // no timing events inside a recording
    if (timed) {
        v0->stage_begin(ST_GATHER);
        (void)hipEventRecord(v0->kev_[0][0], ex.s);
    }
    XCHK(v0, launch_gather_multi(ga.data(), int(ga.size()), ex.ar, ex.s));
    XCHK(v0, launch_bucket_gather_multi(bg.data(), int(bg.size()), ex.ar, ex.s));
    XCHK(v0, launch_gather_color_multi(ca.data(), int(ca.size()), ex.ar, ex.s));
    if (timed) {
        (void)hipEventRecord(v0->kev_[0][1], ex.s);
        v0->kev_used_[0] = true;
        v0->stage_end(ST_GATHER);
    }
    return GVDB_OK;
}

// ----------------------------------------------------------- UpdateApron ----
int Volume::multi_apron(Volume* const* vs, int G, int c, float boundval, const Exec& ex, int* failed) {
    // UpdateApron (gvdb_volume_gvdb.cpp:4418-4457)
    std::vector<NbrArgs> na;
    std::vector<FillArgs> nf;
    std::vector<ApronArgs> aa;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        if (v->apron_ == 0) continue;
        if (c < 0 || c >= 8 || !v->chan_[c].exists) FAILG(g, v->fail(GVDB_ERR_INVALID, "UpdateApron: no such channel"));
        int rc = v->commit_pools();
        if (rc) FAILG(g, rc);
        int bricks = int(v->pool_[0][0].lastEle);
        if (bricks == 0) continue;
        if (ex.cap) ex.cap->note(v);
        if (v->atlas_gen_ != v->topo_gen_)
            FAILG(g, v->fail(GVDB_ERR_STATE, "UpdateApron: call UpdateAtlas after the topology changes"));
        Tree t = v->make_tree();
        if (v->nbr_gen_ != v->topo_gen_) {
            XCHK(v, v->nbr_.ensure(size_t(bricks) * 27 * 4));
            XCHK(v, v->nmask_.ensure(size_t(bricks) * 4));
            nf.push_back(FillArgs{v->nmask_.p, u64(bricks), 0, 0});
            XCHK(v, v->nbase_.ensure(size_t(bricks) * 27 * 8));
            na.push_back(NbrArgs{t, v->nbr_.as<int>(), bricks, v->nmask_.as<uint32_t>(), v->nbase_.as<long long>(),
                                 v->brick_tab()});
            v->nbr_gen_ = v->topo_gen_;
        }
        const bool faces = v->faces_chan_ == c && v->faces_epoch_ == v->atlas_epoch_ && v->faces_topo_ == v->topo_gen_;
        aa.push_back(ApronArgs{t, v->nbase_.as<long long>(), v->nbr_.as<int>(), faces ? v->faces_.p : nullptr,
                               v->chan_[c].data.p, v->chan_[c].type, boundval, bricks});
    }
    if (aa.empty()) return GVDB_OK;
    Volume* v0 = vs[0];
    const bool timed = G == 1 && !ex.cap;
    if (timed) v0->stage_begin(ST_APRON);
    XCHK(v0, launch_fill_multi(nf.data(), int(nf.size()), ex.ar, ex.s));
    XCHK(v0, launch_neighbors_multi(na.data(), int(na.size()), ex.ar, ex.s));
    if (timed) (void)hipEventRecord(v0->kev_[1][0], ex.s);
    XCHK(v0, launch_apron_multi(aa.data(), int(aa.size()), ex.ar, ex.s));
    if (timed) {
        (void)hipEventRecord(v0->kev_[1][1], ex.s);
        v0->kev_used_[1] = true;
        v0->stage_end(ST_APRON);
    }
    return GVDB_OK;
}

// UpdateApron for callers that cannot allocate: the same kernels, with the
// neighbour table in the fixed device array of apron.hip.  Fails instead of
// doing anything that would allocate: pools changed on the host and not yet
// on the device, or more bricks than the array holds.  Every call on one
// device shares that array, so calls that overlap on one device race.
int Volume::update_apron_noalloc(int c, float boundval) {
    if (apron_ == 0) return GVDB_OK;
    if (c < 0 || c >= 8 || !chan_[c].exists) return fail(GVDB_ERR_INVALID, "UpdateApron: no such channel");
    if (dev_stale_)
        return fail(GVDB_ERR_STATE, "UpdateApron (no allocation): the pools changed on the host and are not on the "
                                    "device yet; call UpdateAtlas first");
    const u64 bricks = pool_[0][0].lastEle;
    if (bricks == 0) return GVDB_OK;
    if (atlas_gen_ != topo_gen_) return fail(GVDB_ERR_STATE, "UpdateApron: call UpdateAtlas after the topology changes");
    if (bricks > u64(NOALLOC_MAX_BRICKS))
        return fail(GVDB_ERR_OOM, "UpdateApron (no allocation): " + std::to_string(bricks) + " bricks; the fixed "
                                  "scratch array holds the neighbour table of at most " +
                                  std::to_string(NOALLOC_MAX_BRICKS));
    if (!noalloc_) return fail(GVDB_ERR_HIP, "UpdateApron (no allocation): no scratch array on this device");
    const bool faces = faces_chan_ == c && faces_epoch_ == atlas_epoch_ && faces_topo_ == topo_gen_;
    if (capture_) capture_->note(this);
    stage_begin(ST_APRON);
    if (!capture_) (void)hipEventRecord(kev_[1][0], stream);
    hipError_t e = launch_apron_noalloc(make_tree(), noalloc_, faces ? faces_.p : nullptr, chan_[c].data.p,
                                        chan_[c].type, boundval, int(bricks), stream);
    if (e != hipSuccess) return hip_fail(e, "UpdateApron (no allocation)");
    if (!capture_) {
        (void)hipEventRecord(kev_[1][1], stream);
        kev_used_[1] = true;
    }
    stage_end(ST_APRON);
    return GVDB_OK;
}

// ---------------------------------------------------------------- Render ----
int Volume::multi_render(Volume* const* vs, int G, const int* shading, const int* chan, const int* buf, const Exec& ex,
                         int* failed) {
    // Render (gvdb_volume_gvdb.cpp:4336-4381)
    std::vector<RenderArgs> ra;
    for (int g = 0; g < G; g++) {
        Volume* v = vs[g];
        const int b = buf[g], c = chan[g];
        if (b < 0 || b >= 4 || v->rbuf_[b].w == 0) FAILG(g, v->fail(GVDB_ERR_INVALID, "Render: no such render buffer"));
        if (c < 0 || c >= 8 || !v->chan_[c].exists || v->chan_[c].type == T_UCHAR4)
            FAILG(g, v->fail(GVDB_ERR_INVALID, "Render: no such float or half channel"));
        if (shading[g] != SHADE_VOLUME && shading[g] != SHADE_TRILINEAR)
            FAILG(g, v->fail(GVDB_ERR_UNSUPPORTED, "Render: shading must be SHADE_VOLUME or SHADE_TRILINEAR"));
        if (v->pool_[0][0].lastEle == 0 || v->amap_max_ == 0) FAILG(g, v->fail(GVDB_ERR_STATE, "Render: no topology or atlas"));
        if (ex.cap) ex.cap->note(v);
        if (v->atlas_gen_ != v->topo_gen_)
            FAILG(g, v->fail(GVDB_ERR_STATE, "Render: call UpdateAtlas after the topology changes"));
        int rc = v->commit_pools();
        if (rc) FAILG(g, rc);
        if (!v->tf_committed) {
            rc = v->commit_tf();
            if (rc) FAILG(g, rc);
        }
        rc = v->resolve_bounds();   // the march starts at the object bounds
        if (rc) FAILG(g, rc);
        RenderBuf& r = v->rbuf_[b];
        RenderArgs a;
        a.t = v->make_tree();
        v->scene_block(r.w, r.h, a.s);
        a.tf = v->tf_dev.as<float4>();
        a.atlas = v->chan_[c].data.p;
        a.out = r.data.p;
        a.out_float = r.bpp == 16;
        a.shading = shading[g];
        a.atlas_type = v->chan_[c].type;
        a.tile_w = 8;
        a.clr = nullptr;
        if (v->clr_chan_ != CHAN_UNDEF) {
            // SetColorChannel: the march reads that uchar4 channel's colours
            const int cc = v->clr_chan_;
            if (!v->chan_[cc].exists || v->chan_[cc].type != T_UCHAR4)
                FAILG(g, v->fail(GVDB_ERR_INVALID, "Render: the colour channel must be a uchar4 channel"));
            a.clr = v->chan_[cc].data.as<uint32_t>();
        }
        ra.push_back(a);
    }
    Volume* v0 = vs[0];
    const bool timed = G == 1 && !ex.cap;
    if (timed) {
        v0->stage_begin(ST_RENDER);
        (void)hipEventRecord(v0->kev_[2][0], ex.s);
    }
    XCHK(v0, launch_render_multi(ra.data(), int(ra.size()), ex.ar, ex.s));
    if (timed) {
        (void)hipEventRecord(v0->kev_[2][1], ex.s);
        v0->kev_used_[2] = true;
        v0->stage_end(ST_RENDER);
    }
    return GVDB_OK;
}

}  // namespace gvdb
