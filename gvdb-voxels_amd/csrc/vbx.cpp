// vbx.cpp -- SaveVBX / LoadVBX: GVDB's VBX file, byte for byte.
//
// Layout (src/gvdb_volume_gvdb.cpp:1626-1768 SaveVBX, :507-685 LoadVBX;
// GVDB_FILESPEC.txt), little-endian, no padding:
//   u8 major, u8 minor                      (1, 11)
//   vec3f pretrans, angs, scale, trans      (versions 1.11 and later)
//   i32 grids                               (u8 bitmask flag from version 2)
//   u64 grid offsets[grids]
// per grid:
//   char name[256]; u8 dtype 'f'; u8 components 1; u8 compression 0;
//   vec3f voxel size (1,1,1); i32 bricks; vec3i brick dims; i32 apron;
//   i32 channels; u64 atlas bytes (channel 0); u8 topology 2; i32 reuse 0;
//   u8 layout 0; vec3i atlas brick counts; vec3i atlas resolution;
//   i32 levels; u64 root;
//   per level: i32 log2dim, i32 res, vec3i range, i32 count0, i32 width0,
//              i32 count1, i32 width1;
//   pool 0 of every level (count0 rows of width0 bytes), then pool 1;
//   per channel: i32 type, i32 bytes per voxel, the 3D atlas in z slices
//              (x fastest).
// The atlas is brick-major on the device; it is converted to and from the
// 3D layout on the host.
#include "volume.h"
#include "../../include/gvdb_amd.h"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <vector>
#include <memory>

namespace gvdb {

#define VBX_HIPCHK(x)                                      \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return hip_fail(e_, #x);     \
    } while (0)

namespace {
struct File {
    FILE* f = nullptr;
    bool ok = true;
    ~File() { if (f) fclose(f); }
    template <typename T> void put(const T& v) { ok = ok && fwrite(&v, sizeof(T), 1, f) == 1; }
    void put_bytes(const void* p, size_t n) { ok = ok && (n == 0 || fwrite(p, 1, n, f) == n); }
    template <typename T> T get() {
        T v{};
        ok = ok && fread(&v, sizeof(T), 1, f) == 1;
        return v;
    }
    void get_bytes(void* p, size_t n) { ok = ok && (n == 0 || fread(p, 1, n, f) == n); }
};
const uint8_t VBX_MAJOR = 1, VBX_MINOR = 11;   // src/gvdb_volume_gvdb.cpp:31-32
}  // namespace

int Volume::save_vbx(const char* path) {
    if (levels_ == 0) return fail(GVDB_ERR_STATE, "SaveVBX before Configure");
    if (!path) return fail(GVDB_ERR_INVALID, "SaveVBX: null path");
    int rc = sync_host();
    if (rc) return rc;
    File fp;
    fp.f = fopen(path, "wb");
    if (!fp.f) return fail(GVDB_ERR_IO, std::string("SaveVBX: cannot open ") + path);
    int nchan = 0;
    while (nchan < 8 && chan_[nchan].exists) nchan++;
    const int br = brickwid_ + 2 * apron_;
    const int axiscnt[3] = {nchan ? atlas_cnt_[0] : 0, nchan ? atlas_cnt_[1] : 0, nchan ? atlas_cnt_[2] : 0};
    const int axisres[3] = {axiscnt[0] * br, axiscnt[1] * br, axiscnt[2] * br};
    const u64 vox = u64(axisres[0]) * axisres[1] * axisres[2];
    const u64 atlas_sz = nchan ? vox * u64(chan_[0].elem()) : 0;

    fp.put(VBX_MAJOR);
    fp.put(VBX_MINOR);
    for (int a = 0; a < 3; a++) fp.put(xf_pretrans_[a]);
    for (int a = 0; a < 3; a++) fp.put(xf_angs_[a]);
    for (int a = 0; a < 3; a++) fp.put(xf_scale_[a]);
    for (int a = 0; a < 3; a++) fp.put(xf_trans_[a]);
    fp.put(int32_t(1));
    const long table = ftell(fp.f);
    fp.put(u64(0));
    const u64 grid_off = u64(ftell(fp.f));
    char name[256];
    memset(name, 0, sizeof(name));
    fp.put_bytes(name, sizeof(name));
    fp.put(uint8_t('f'));
    fp.put(uint8_t(1));
    fp.put(uint8_t(0));
    for (int a = 0; a < 3; a++) fp.put(1.0f);
    fp.put(int32_t(pool_[0][0].lastEle));
    for (int a = 0; a < 3; a++) fp.put(int32_t(res_[0]));
    fp.put(int32_t(apron_));
    fp.put(int32_t(nchan));
    fp.put(atlas_sz);
    fp.put(uint8_t(2));
    fp.put(int32_t(0));
    fp.put(uint8_t(0));
    for (int a = 0; a < 3; a++) fp.put(int32_t(axiscnt[a]));
    for (int a = 0; a < 3; a++) fp.put(int32_t(axisres[a]));
    fp.put(int32_t(levels_));
    fp.put(u64(root_));
    for (int n = 0; n < levels_; n++) {
        fp.put(int32_t(logdim_[n]));
        fp.put(int32_t(res_[n]));
        for (int a = 0; a < 3; a++) fp.put(int32_t(range_[n]));
        fp.put(int32_t(pool_[0][n].lastEle));
        fp.put(int32_t(pool_[0][n].stride));
        fp.put(int32_t(n > 0 ? pool_[1][n].lastEle : 0));
        fp.put(int32_t(n > 0 ? pool_[1][n].stride : 0));
    }
    for (int g = 0; g < 2; g++)
        for (int n = 0; n < levels_; n++) {
            const Pool& p = pool_[g][n];
            if (g == 1 && n == 0) continue;
            fp.put_bytes(p.host.data(), size_t(p.lastEle * p.stride));
        }
    for (int c = 0; c < nchan; c++) {
        fp.put(int32_t(chan_[c].type));
        fp.put(int32_t(chan_[c].elem()));
        std::vector<uint8_t> a3(size_t(vox) * chan_[c].elem());
        rc = read_atlas(c, a3.data());
        if (rc) return rc;
        fp.put_bytes(a3.data(), a3.size());
    }
    fseek(fp.f, table, SEEK_SET);
    fp.put(grid_off);
    if (!fp.ok) return fail(GVDB_ERR_IO, std::string("SaveVBX: write failed on ") + path);
    return GVDB_OK;
}

int Volume::load_vbx(const char* path, int force_major, int force_minor) {
    if (!path) return fail(GVDB_ERR_INVALID, "LoadVBX: null path");
    File fp;
    fp.f = fopen(path, "rb");
    if (!fp.f) return fail(GVDB_ERR_IO, std::string("LoadVBX: unable to open ") + path);
    uint8_t major = fp.get<uint8_t>(), minor = fp.get<uint8_t>();
    if (force_major > 0 || force_minor > 0) { major = uint8_t(force_major); minor = uint8_t(force_minor); }
    float xf[12] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0};   // pretrans, angs, scale, trans
    if ((major == 1 && minor >= 11) || major > 1)
        for (float& v : xf) v = fp.get<float>();
    const int32_t grids = fp.get<int32_t>();
    uint8_t masks = 0;
    if (major >= 2) masks = fp.get<uint8_t>();
    else if (major == 1 && minor == 0) masks = 1;
    if (!fp.ok || grids < 1 || grids > (1 << 20)) return fail(GVDB_ERR_IO, "LoadVBX: truncated or invalid header");
    if (masks) return fail(GVDB_ERR_UNSUPPORTED, "LoadVBX: nodes with bitmasks are not supported");
    for (int g = 0; g < grids; g++) (void)fp.get<u64>();
    // the reference reads the grids one after another into the same volume
    for (int g = 0; g < grids; g++) {
        char name[256];
        fp.get_bytes(name, sizeof(name));
        const uint8_t dtype = fp.get<uint8_t>(), comps = fp.get<uint8_t>(), compress = fp.get<uint8_t>();
        for (int a = 0; a < 3; a++) (void)fp.get<float>();
        (void)fp.get<int32_t>();   // bricks (the node pool says how many)
        int32_t leafdim[3];
        for (int a = 0; a < 3; a++) leafdim[a] = fp.get<int32_t>();
        const int32_t apron = fp.get<int32_t>();
        const int32_t nchan = fp.get<int32_t>();
        (void)fp.get<u64>();       // atlas bytes
        const uint8_t topo = fp.get<uint8_t>();
        (void)fp.get<int32_t>();   // topology reuse
        const uint8_t layout = fp.get<uint8_t>();
        int32_t axiscnt[3], axisres[3];
        for (int a = 0; a < 3; a++) axiscnt[a] = fp.get<int32_t>();
        for (int a = 0; a < 3; a++) axisres[a] = fp.get<int32_t>();
        const int32_t levels = fp.get<int32_t>();
        const u64 root = fp.get<u64>();
        if (!fp.ok) return fail(GVDB_ERR_IO, "LoadVBX: truncated grid header");
        if (dtype != 'f' || comps != 1 || compress != 0 || topo != 2 || layout != 0)
            return fail(GVDB_ERR_UNSUPPORTED, "LoadVBX: only uncompressed scalar float grids with GVDB topology "
                                              "in atlas layout");
        if (levels < 2 || levels > MAXLEV || nchan < 0 || nchan > 8 || apron < 0 || apron > 1)
            return fail(GVDB_ERR_IO, "LoadVBX: invalid grid header");
        int ld[MAXLEV], cnt0[MAXLEV], w0[MAXLEV], cnt1[MAXLEV], w1[MAXLEV];
        for (int n = 0; n < levels; n++) {
            ld[n] = fp.get<int32_t>();
            (void)fp.get<int32_t>();   // res
            for (int a = 0; a < 3; a++) (void)fp.get<int32_t>();   // range
            cnt0[n] = fp.get<int32_t>();
            w0[n] = fp.get<int32_t>();
            cnt1[n] = fp.get<int32_t>();
            w1[n] = fp.get<int32_t>();
        }
        if (!fp.ok) return fail(GVDB_ERR_IO, "LoadVBX: truncated topology header");
        if (w0[0] != int(sizeof(Node)))
            return fail(GVDB_ERR_IO, "LoadVBX: VBX file contains nodes incompatible with this library");
        int rc = configure(levels, ld, cnt0);
        if (rc) return rc;
        if (leafdim[0] != res_[0] || leafdim[1] != res_[0] || leafdim[2] != res_[0])
            return fail(GVDB_ERR_IO, "LoadVBX: brick dimensions differ from the level-0 resolution");
        for (int grp = 0; grp < 2; grp++)
            for (int n = 0; n < levels; n++) {
                const int cnt = grp ? cnt1[n] : cnt0[n], wid = grp ? w1[n] : w0[n];
                if (grp == 1 && n == 0) {
                    if (cnt != 0 && wid != 0) return fail(GVDB_ERR_IO, "LoadVBX: level 0 has no child lists");
                    continue;
                }
                Pool& p = pool_[grp][n];
                if (cnt < 0 || u64(wid) != p.stride) return fail(GVDB_ERR_IO, "LoadVBX: pool width differs");
                while (p.max < u64(cnt)) p.max *= 2;
                p.host.assign(size_t(p.max * p.stride), 0);
                fp.get_bytes(p.host.data(), size_t(cnt) * wid);
                p.lastEle = p.usedNum = u64(cnt);
            }
        if (!fp.ok) return fail(GVDB_ERR_IO, "LoadVBX: truncated pools");
        // The kernels walk the pools without bounds checks, so every reference
        // a file holds must name a node or child list that exists: child list
        // ids, the child slots of each list, and parents.
        for (int n = 0; n < levels; n++) {
            const Pool& p0 = pool_[0][n];
            for (u64 i = 0; i < p0.lastEle; i++) {
                const Node* nd = reinterpret_cast<const Node*>(p0.host.data() + i * p0.stride);
                const u64 par = nd->mParent, cl = nd->mChildList;
                if (int(nd->mLev) != n) return fail(GVDB_ERR_IO, "LoadVBX: node level differs from its pool");
                // UpdateAtlas rewrites the corners of used leaves; an unused
                // one keeps its own, which must stay inside the atlas
                if (n == 0 && !nd->mFlags)
                    for (int a = 0; a < 3; a++)
                        if (nd->mValue[a] < -1 || nd->mValue[a] >= axisres[a])
                            return fail(GVDB_ERR_IO, "LoadVBX: unused brick's atlas corner outside the atlas");
                if (par != ID_UNDEF64 && par != ID_UNDEFL &&
                    (n + 1 >= levels || (par & 0xFF) != 0 || int((par >> 8) & 0xFF) != n + 1 ||
                     (par >> 16) >= pool_[0][n + 1].lastEle))
                    return fail(GVDB_ERR_IO, "LoadVBX: parent reference outside the pools");
                if (n == 0 || cl == ID_UNDEFL) continue;
                if ((cl & 0xFF) != 1 || int((cl >> 8) & 0xFF) != n || (cl >> 16) >= pool_[1][n].lastEle)
                    return fail(GVDB_ERR_IO, "LoadVBX: child list reference outside the pools");
            }
            if (n == 0) continue;
            const Pool& p1 = pool_[1][n];
            const u64 slots = voxcnt(n);
            for (u64 i = 0; i < p1.lastEle * slots; i++) {
                const u64 c = reinterpret_cast<const u64*>(p1.host.data())[i];
                if (c != ID_UNDEF64 && ((c & 0xFF) != 0 || int((c >> 8) & 0xFF) != n - 1 ||
                                        (c >> 16) >= pool_[0][n - 1].lastEle))
                    return fail(GVDB_ERR_IO, "LoadVBX: child slot outside the pools");
            }
        }
        if (root != ID_UNDEFL && ((root & 0xFF) != 0 || int((root >> 8) & 0xFF) >= levels ||
                                  (root >> 16) >= pool_[0][(root >> 8) & 0xFF].lastEle))
            return fail(GVDB_ERR_IO, "LoadVBX: root reference outside the pools");
        root_ = root;
        host_valid_ = true;
        dev_stale_ = true;
        compute_top_lev();
        {
            // the file's grid transform, as SetTransform sets it
            const float pt[3] = {xf[0], xf[1], xf[2]}, an[3] = {xf[3], xf[4], xf[5]}, sc[3] = {xf[6], xf[7], xf[8]},
                        tr[3] = {xf[9], xf[10], xf[11]};
            rc = set_transform(pt, sc, an, tr);
            if (rc) return rc;
        }
        rc = finish_topology(true, true);
        if (rc) return rc;
        // DestroyChannels, then one AddChannel per stored channel
        for (auto& ch : chan_) {
            ch.data.release();
            ch.exists = false;
        }
        const int br = brickwid_ + 2 * apron;
        for (int c = 0; c < nchan; c++) {
            const int32_t type = fp.get<int32_t>(), stride = fp.get<int32_t>();
            if (!fp.ok) return fail(GVDB_ERR_IO, "LoadVBX: truncated channel header");
            if (axisres[0] != axiscnt[0] * br || axisres[1] != axiscnt[1] * br || axisres[2] != axiscnt[2] * br)
                return fail(GVDB_ERR_IO, "LoadVBX: atlas resolution differs from brick count x padded brick");
            rc = add_channel(c, type, apron, 0, 0, axiscnt[0], axiscnt[1], axiscnt[2]);
            if (rc) return rc;
            if (stride != chan_[c].elem()) return fail(GVDB_ERR_IO, "LoadVBX: voxel size differs from the channel type");
            const u64 vox = u64(axisres[0]) * axisres[1] * axisres[2];
            std::vector<uint8_t> a3(size_t(vox) * stride);
            fp.get_bytes(a3.data(), a3.size());
            if (!fp.ok) return fail(GVDB_ERR_IO, "LoadVBX: truncated atlas");
            rc = write_atlas(c, a3.data());
            if (rc) return rc;
        }
        if (nchan > 0) {
            rc = update_atlas();
            if (rc) return rc;
        }
    }
    return GVDB_OK;
}

void Volume::get_transform(float out12[12]) const {
    for (int a = 0; a < 3; a++) {
        out12[a] = xf_pretrans_[a];
        out12[3 + a] = xf_angs_[a];
        out12[6 + a] = xf_scale_[a];
        out12[9 + a] = xf_trans_[a];
    }
}

bool Volume::identity_transform() const {
    for (int a = 0; a < 3; a++)
        if (xf_pretrans_[a] != 0 || xf_angs_[a] != 0 || xf_scale_[a] != 1 || xf_trans_[a] != 0) return false;
    return true;
}

int Volume::write_atlas(int c, const void* src) {
    // GVDB's 3D atlas layout (x fastest) -> device brick-major storage
    if (c < 0 || c >= 8 || !chan_[c].exists || !src) return fail(GVDB_ERR_INVALID, "write_atlas: no such channel");
    Channel& ch = chan_[c];
    const int br = brickwid_ + 2 * apron_;
    const u64 slots = atlas_max_;
    const size_t bv = size_t(br) * br * br;
    const int e = ch.elem();
    std::vector<uint8_t> tmp(size_t(slots) * bv * e);
    const int rx = atlas_cnt_[0] * br, ry = atlas_cnt_[1] * br;
    const uint8_t* in = reinterpret_cast<const uint8_t*>(src);
    for (u64 s = 0; s < slots; s++) {
        const int sx = int(s % atlas_cnt_[0]), sy = int((s / atlas_cnt_[0]) % atlas_cnt_[1]);
        const int sz = int(s / (u64(atlas_cnt_[0]) * atlas_cnt_[1]));
        for (int z = 0; z < br; z++)
            for (int y = 0; y < br; y++) {
                const size_t d = (s * bv + size_t(z * br + y) * br) * e;
                const size_t o = ((size_t(sz * br + z) * ry + (sy * br + y)) * rx + sx * br) * e;
                memcpy(tmp.data() + d, in + o, size_t(br) * e);
            }
    }
    VBX_HIPCHK(hipMemcpyAsync(ch.data.p, tmp.data(), tmp.size(), hipMemcpyHostToDevice, stream));
    VBX_HIPCHK(hipStreamSynchronize(stream));
    atlas_epoch_++;
    return GVDB_OK;
}

}  // namespace gvdb
