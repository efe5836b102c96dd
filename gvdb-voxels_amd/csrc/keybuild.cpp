// keybuild.cpp -- Volume::keyed_build: RebuildTopology's full build over
// sorted brick keys (kernels in keysort.hip), taken when the points' level-0
// brick box holds more than 2^34 bricks.  Work and memory follow the number
// of points and of dilated bricks, not the box: per level, the cells of the
// own bricks (A) and of the dilated bricks (B, which contains A) are sorted
// and made unique, the parents of level L's cells give level L+1's, and each
// level's nodes are A's cells then B's cells outside A, in key order, as the
// bitmap build orders them (ActivateBricksGPU then ActivateExtraBricksGPU:
// src/gvdb_volume_gvdb.cpp:1181-1331, :935-1018).
#include "volume.h"
#include "../../include/gvdb_amd.h"
#include <algorithm>
#include <cstring>

namespace gvdb {

#define KCHK(x)                                                  \
    do {                                                         \
        hipError_t e_ = (x);                                     \
        if (e_ != hipSuccess) return hip_fail(e_, #x);           \
    } while (0)

// keys and dilated bricks per build: past this the sort's 32-bit offsets
static const u64 KEYS_MAX = u64(1) << 31;

static int key_bits(u64 nbits) {
    int b = 1;
    while (b < 64 && (u64(1) << b) <= nbits) b++;
    return b;   // enough for 0..nbits (the sentinel is nbits)
}

namespace {
// per-level lists owned by one build
struct Lists {
    std::vector<DevBuf> bufs;
    DevBuf& add() {
        bufs.emplace_back();
        return bufs.back();
    }
    ~Lists() {
        for (DevBuf& b : bufs) b.release();
    }
};
}  // namespace

// Sorts keys[0, cnt) (`bits` significant) and keeps the distinct ones other
// than `sentinel`, in order, in `out`; *count gets their number.  keys, kb_tmp_
// and the flag buffers are scratch; the call waits for the count.
int Volume::kb_sort_unique(u64* keys, u64 cnt, int bits, u64 sentinel, DevBuf& out, uint32_t* count, const Exec& ex) {
    const hipStream_t s = ex.s;
    *count = 0;
    if (cnt == 0) return GVDB_OK;
    KCHK(kb_tmp_.ensure(size_t(cnt) * 8));
    const u64 hw = radix_hist_words(cnt);
    KCHK(kb_hist_.ensure(size_t(hw + scan_tmp_words(hw) + 64) * 4));
    uint32_t* hist = kb_hist_.as<uint32_t>();
    u64* sorted = nullptr;
    KCHK(radix_sort_u64(keys, kb_tmp_.as<u64>(), cnt, bits, hist, hist + hw + 32, s, &sorted));
    KCHK(kb_flags_.ensure(size_t(cnt) * 4));
    KCHK(kb_pos_.ensure(size_t(cnt) * 4));
    KCHK(scan_tmp_.ensure(scan_tmp_words(cnt) * 4 + 256));
    KCHK(launch_flag_unique(sorted, cnt, sentinel, kb_flags_.as<uint32_t>(), s));
    KCHK(launch_scan_u32(kb_flags_.as<uint32_t>(), nullptr, 0, kb_pos_.as<uint32_t>(), cnt, ex.dsmall + SM_TOTAL,
                         scan_tmp_.as<uint32_t>(), s));
    KCHK(hipMemcpyAsync(ex.hsmall, ex.dsmall, SMALL_WORDS * 4, hipMemcpyDeviceToHost, s));
    KCHK(hipStreamSynchronize(s));
    *count = ex.hsmall[SM_TOTAL];
    KCHK(out.ensure(size_t(std::max<uint32_t>(*count, 1)) * 8));
    KCHK(launch_compact(sorted, cnt, kb_flags_.as<uint32_t>(), kb_pos_.as<uint32_t>(), out.as<u64>(), s));
    return GVDB_OK;
}

void Volume::kb_release() {
    for (DevBuf* b : {&kb_keys_, &kb_keysB_, &kb_cnt_, &kb_off_, &kb_tmp_, &kb_hist_, &kb_flags_, &kb_pos_, &kb_list_})
        b->release();
}

int Volume::keyed_build(int n, float radius, const float orig[3], int D, const int rootPos[3], const BitDomain* dom,
                        const Exec& ex) {
    const hipStream_t s = ex.s;
    const bool dilate = radius >= 1.0f;
    uint32_t* small = ex.dsmall;
    const uint32_t* hsmall = ex.hsmall;
    auto sync = [&]() -> int {
        KCHK(hipMemcpyAsync(ex.hsmall, ex.dsmall, SMALL_WORDS * 4, hipMemcpyDeviceToHost, s));
        KCHK(hipStreamSynchronize(s));
        return GVDB_OK;
    };
    auto read64 = [&](int w) {
        u64 v = 0;
        memcpy(&v, hsmall + w, 8);
        return v;
    };
    // out = the keys of b (sorted, distinct) absent from a (sorted), in order
    auto difference = [&](const DevBuf& b, uint32_t nb, const DevBuf& a, uint32_t na, DevBuf& out,
                          uint32_t* count) -> int {
        *count = 0;
        if (nb == 0) return GVDB_OK;
        KCHK(kb_flags_.ensure(size_t(nb) * 4));
        KCHK(kb_pos_.ensure(size_t(nb) * 4));
        KCHK(scan_tmp_.ensure(scan_tmp_words(nb) * 4 + 256));
        KCHK(launch_flag_absent(b.as<u64>(), nb, a.as<u64>(), na, kb_flags_.as<uint32_t>(), s));
        KCHK(launch_scan_u32(kb_flags_.as<uint32_t>(), nullptr, 0, kb_pos_.as<uint32_t>(), nb, small + SM_TOTAL,
                             scan_tmp_.as<uint32_t>(), s));
        int rc = sync();
        if (rc) return rc;
        *count = hsmall[SM_TOTAL];
        KCHK(out.ensure(size_t(std::max<uint32_t>(*count, 1)) * 8));
        KCHK(launch_compact(b.as<u64>(), nb, kb_flags_.as<uint32_t>(), kb_pos_.as<uint32_t>(), out.as<u64>(), s));
        return GVDB_OK;
    };

    // 1. level-0 keys: one per point (its own brick) and its dilation box
    const BitDomain& d0 = dom[0];
    KCHK(kb_keys_.ensure(size_t(std::max(n, 1)) * 8));
    KCHK(kb_cnt_.ensure(size_t(std::max(n, 1)) * 4));
    KCHK(kb_off_.ensure(size_t(std::max(n, 1)) * 4));
    PointSrc P = pts_;
    P.n = n;
    KeyGenArgs kg{P, {orig[0], orig[1], orig[2]}, radius, dilate ? 1 : 0, int(range_[0]), 1.0f / float(range_[0]),
                  d0, d0.nbits, kb_keys_.as<u64>(), kb_cnt_.as<uint32_t>(), kb_off_.as<uint32_t>(), nullptr};
    KCHK(launch_key_gen(kg, false, s));
    u64 totalB = 0;
    if (dilate) {
        KCHK(launch_sum_u32(kb_cnt_.as<uint32_t>(), u64(n), reinterpret_cast<u64*>(small + SM_LEN), s));
        KCHK(scan_tmp_.ensure(scan_tmp_words(u64(n)) * 4 + 256));
        KCHK(launch_scan_u32(kb_cnt_.as<uint32_t>(), nullptr, 0, kb_off_.as<uint32_t>(), u64(n), nullptr,
                             scan_tmp_.as<uint32_t>(), s));
        int rc = sync();
        if (rc) return rc;
        totalB = read64(SM_LEN);
        if (totalB > KEYS_MAX)
            return fail(GVDB_ERR_RANGE, "RebuildTopology: more than 2^31 dilated bricks over the points (" +
                                            std::to_string(totalB) + ")");
        KCHK(kb_keysB_.ensure(size_t(std::max<u64>(totalB, 1)) * 8));
        kg.keysB = kb_keysB_.as<u64>();
        KCHK(launch_key_gen(kg, true, s));
    }

    // 2. per level: A and B lists (sorted, distinct), E = B \ A
    Lists lists;
    lists.bufs.reserve(3 * size_t(D));   // the pointers below stay valid
    std::vector<DevBuf*> LA(D), LB(D), LE(D);
    std::vector<uint32_t> na(D, 0), nb(D, 0), ne(D, 0);
    int rc = 0;
    for (int L = 0; L < D; L++) {
        LA[L] = &lists.add();
        LB[L] = &lists.add();
        LE[L] = &lists.add();
    }
    rc = kb_sort_unique(kb_keys_.as<u64>(), u64(n), key_bits(d0.nbits), d0.nbits, *LA[0], &na[0], ex);
    if (rc) return rc;
    if (dilate) {
        rc = kb_sort_unique(kb_keysB_.as<u64>(), totalB, key_bits(d0.nbits), d0.nbits, *LB[0], &nb[0], ex);
        if (rc) return rc;
    }
    for (int L = 1; L < D; L++) {
        const int ld = logdim_[L];
        // parents of the level below, into the reusable key buffers, then sorted
        KCHK(kb_keys_.ensure(size_t(std::max<uint32_t>(na[L - 1], 1)) * 8));
        KCHK(launch_parent_keys(LA[L - 1]->as<u64>(), na[L - 1], dom[L - 1], ld, dom[L], kb_keys_.as<u64>(), s));
        rc = kb_sort_unique(kb_keys_.as<u64>(), na[L - 1], key_bits(dom[L].nbits), dom[L].nbits, *LA[L], &na[L], ex);
        if (rc) return rc;
        if (dilate) {
            KCHK(kb_keysB_.ensure(size_t(std::max<uint32_t>(nb[L - 1], 1)) * 8));
            KCHK(launch_parent_keys(LB[L - 1]->as<u64>(), nb[L - 1], dom[L - 1], ld, dom[L], kb_keysB_.as<u64>(), s));
            rc = kb_sort_unique(kb_keysB_.as<u64>(), nb[L - 1], key_bits(dom[L].nbits), dom[L].nbits, *LB[L], &nb[L], ex);
            if (rc) return rc;
        }
    }
    for (int L = 0; L < D && dilate; L++) {
        rc = difference(*LB[L], nb[L], *LA[L], na[L], *LE[L], &ne[L]);
        if (rc) return rc;
    }

    // 3. pools, nodes and links (as the bitmap build's last steps)
    std::vector<KeyLevel> KL(D);
    for (int L = 0; L < D; L++) {
        KL[L] = KeyLevel{dom[L], LA[L]->as<u64>(), na[L], LE[L]->as<u64>(), ne[L], L, int(range_[L])};
        const u64 cnt = u64(na[L]) + ne[L];
        pool_[0][L].lastEle = pool_[0][L].usedNum = cnt;
        rc = ensure_pool_capacity(0, L, std::max<u64>(cnt, 1));
        if (rc) return rc;
        if (L > 0) {
            pool_[1][L].lastEle = pool_[1][L].usedNum = cnt;
            rc = ensure_pool_capacity(1, L, std::max<u64>(cnt, 1));
            if (rc) return rc;
        }
    }
    pool_[0][D].lastEle = pool_[0][D].usedNum = 1;
    pool_[1][D].lastEle = pool_[1][D].usedNum = 1;
    rc = ensure_pool_capacity(0, D, 1);
    if (rc) return rc;
    rc = ensure_pool_capacity(1, D, 1);
    if (rc) return rc;
    compute_top_lev();
    RootArgs r;
    r.root = pool_[0][D].dev.as<Node>();
    for (int a = 0; a < 3; a++) r.pos[a] = rootPos[a];
    r.lev = D;
    KCHK(launch_write_root_multi(&r, 1, nullptr, s));
    std::vector<FillArgs> cf;
    cf.push_back(FillArgs{pool_[1][D].dev.p, voxcnt(D), ID_UNDEF64, 1});
    for (int L = 1; L < D; L++) cf.push_back(FillArgs{pool_[1][L].dev.p, pool_[1][L].lastEle * voxcnt(L), ID_UNDEF64, 1});
    for (const FillArgs& f : cf) KCHK(launch_fill_multi(&f, 1, nullptr, s));
    for (int L = 0; L < D; L++) KCHK(launch_emit_list(KL[L], pool_[0][L].dev.as<Node>(), L > 0 ? 1 : 0, s));
    // gvdbLinkBricks walks from the top node; levels above it stay unlinked
    for (int L = 0; L < D; L++) {
        if (L + 1 > top_lev_) continue;
        const bool toRoot = (L + 1 == D);
        KCHK(launch_link_list(KL[L], toRoot ? KL[L] : KL[L + 1], toRoot ? 1 : 0, logdim_[L + 1],
                              pool_[0][L].dev.as<Node>(), pool_[1][L + 1].dev.as<u64>(), voxcnt(L + 1), s));
    }
    // the lists and the key buffers are freed here: the launches above must be done
    KCHK(hipStreamSynchronize(s));
    kb_release();
    root_ = Elem(0, D, 0);
    rebuild_topo_ = false;
    host_valid_ = false;
    dev_stale_ = false;
    topo_gen_++;
    return GVDB_OK;
}

}  // namespace gvdb
