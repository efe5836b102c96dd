// render.hip -- hierarchical DDA ray march over the sparse tree.
//
// Follows rayCast (kernels/cuda_gvdb_raycast.cuh:543-611 of the reference)
// with the HDDA stepping of kernels/cuda_gvdb_dda.cuh:38-91, the
// emission-absorption brick march rayDeepBrick (:485-533) and the trilinear
// isosurface march raySurfaceTrilinearBrick (:281-300) with getGradient
// (:132-141) and Phong + shadow ray (kernels/cuda_gvdb_module.cu:38-57).
// Samples are exact software trilinear reads of the brick-major atlas; the
// reference's texture unit is replaced because its filter weights are
// quantised.  One 64-thread workgroup renders one pixel tile (8x8 by
// default; 16x4, 32x2 or 64x1 when the autotuner finds them faster).
#include "kernels.h"
#include "tree_device.h"
#include <hip/hip_fp16.h>

// Contraction is off in this file: the walk performs the reference's FP32
// operations one by one (no fused multiply-adds), so every ray takes the
// steps of a plain IEEE single-precision evaluation of rayCast, also where a
// step sits on a rounding edge (far from the origin the 1e-3 epsilon is a
// few ulps).
#pragma clang fp contract(off)

namespace gvdb {

struct f3 { float x, y, z; };
__device__ inline f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
__device__ inline f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 norm3(f3 a) { float il = 1.0f / sqrtf(dot3(a, a)); return a * il; }

template <typename AT>
__device__ inline float ldv(const AT* a, u64 o);
template <>
__device__ inline float ldv<float>(const float* a, u64 o) { return a[o]; }
template <>
__device__ inline float ldv<__half>(const __half* a, u64 o) { return __half2float(a[o]); }

// Trilinear sample at atlas coordinate (x, y, z) = p + mValue, the sum formed
// in FP32 as the reference's tex3D call forms it (far from the atlas origin
// the position carries that sum's rounding), read from the padded brick
// whose first voxel is atlas voxel (cx, cy, cz) and whose data starts at
// `b`: an unnormalised linear fetch (texel centres at integer + 0.5) with
// exact weights, blended as a + f (b - a) along x, then y, then z.  Split
// into the fetch (four x-adjacent pairs, one 8-byte load each for FP32) and
// the blend, so the loads of several samples can be in flight together.
// A march sample (p in [0, res)) has its lower taps at padded indices 0..8;
// indices are clamped into the brick, so a fetch at any position stays
// inside it.
struct Taps {
    float c[8];   // c000 c100 c010 c110 c001 c101 c011 c111
    float fx, fy, fz;
};

template <typename AT>
__device__ inline void ld_pair(const AT* p, float& a, float& b);
template <>
__device__ inline void ld_pair<float>(const float* p, float& a, float& b) {
    float2 v;
    __builtin_memcpy(&v, p, sizeof(v));
    a = v.x;
    b = v.y;
}
template <>
__device__ inline void ld_pair<__half>(const __half* p, float& a, float& b) {
    __half2 v;
    __builtin_memcpy(&v, p, sizeof(v));
    a = __low2float(v);
    b = __high2float(v);
}

struct Corner {
    int c[3];      // atlas voxel of the padded brick's first voxel (mValue - apron)
    float o[3];    // mValue as floats: atlas position = p + o
};

template <typename AT>
__device__ inline void tri_fetch(const AT* b, int br, const Corner& K, float x, float y, float z, Taps& T) {
    const float ux = x - 0.5f, uy = y - 0.5f, uz = z - 0.5f;
    const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    const int ix = min(max(int(fx0) - K.c[0], 0), br - 2), iy = min(max(int(fy0) - K.c[1], 0), br - 2),
              iz = min(max(int(fz0) - K.c[2], 0), br - 2);
    T.fx = ux - fx0;
    T.fy = uy - fy0;
    T.fz = uz - fz0;
    const AT* o = b + ((iz * br + iy) * br + ix);
    const int sy = br, sz = br * br;
    ld_pair<AT>(o, T.c[0], T.c[1]);
    ld_pair<AT>(o + sy, T.c[2], T.c[3]);
    ld_pair<AT>(o + sz, T.c[4], T.c[5]);
    ld_pair<AT>(o + sz + sy, T.c[6], T.c[7]);
}

__device__ inline float tri_blend(const Taps& T) {
    const float* c = T.c;
    float c00 = c[0] + T.fx * (c[1] - c[0]), c10 = c[2] + T.fx * (c[3] - c[2]);
    float c01 = c[4] + T.fx * (c[5] - c[4]), c11 = c[6] + T.fx * (c[7] - c[6]);
    float c0 = c00 + T.fy * (c10 - c00), c1 = c01 + T.fy * (c11 - c01);
    return c0 + T.fz * (c1 - c0);
}

// The same sample where the lower tap may be the brick's last padded voxel
// (the gradient's tap half a voxel past a march sample): that tap's weight
// is 1 and its upper neighbour (outside the brick) is weighted 0, so the
// fetch takes the pair below and moves the upper value down.
template <typename AT>
__device__ inline float tri_edge(const AT* b, int br, const Corner& K, float x, float y, float z) {
    Taps T;
    tri_fetch<AT>(b, br, K, x, y, z, T);
    float* c = T.c;
    if (int(floorf(x - 0.5f)) - K.c[0] == br - 1)
        for (int k = 0; k < 8; k += 2) c[k] = c[k + 1];
    if (int(floorf(y - 0.5f)) - K.c[1] == br - 1) {
        c[0] = c[2]; c[1] = c[3]; c[4] = c[6]; c[5] = c[7];
    }
    if (int(floorf(z - 0.5f)) - K.c[2] == br - 1)
        for (int k = 0; k < 4; k++) c[k] = c[k + 4];
    return tri_blend(T);
}

// getColorF (kernels/cuda_gvdb_raycast.cuh:206-209): the colour channel's
// voxel at the integer part of the atlas position, bytes / 255
__device__ inline float4 color_at(const uint32_t* cb, int br, const Corner& K, float x, float y, float z) {
    const int ix = min(max(int(x) - K.c[0], 0), br - 1), iy = min(max(int(y) - K.c[1], 0), br - 1),
              iz = min(max(int(z) - K.c[2], 0), br - 1);
    const uint32_t c = cb[(iz * br + iy) * br + ix];
    return make_float4(float(c & 0xFFu) / 255.0f, float((c >> 8) & 0xFFu) / 255.0f, float((c >> 16) & 0xFFu) / 255.0f,
                       float(c >> 24) / 255.0f);
}

template <typename AT>
__device__ inline Corner brick_corner(const Tree& t, const Node* n) {
    Corner K;
    for (int a = 0; a < 3; a++) {
        K.c[a] = n->mValue[a] - t.apron;
        K.o[a] = float(n->mValue[a]);
    }
    return K;
}

struct Hdda {
    f3 pos, dir, tDel, t, tSide;
    int ps[3], p[3], m[3];
    __device__ void prepare(f3 vmin, float vdel) {
        tDel = mk3(fabsf(vdel / dir.x), fabsf(vdel / dir.y), fabsf(vdel / dir.z));
        float pf[3] = {(pos.x + t.x * dir.x - vmin.x) / vdel, (pos.y + t.x * dir.y - vmin.y) / vdel,
                       (pos.z + t.x * dir.z - vmin.z) / vdel};
        float td[3] = {tDel.x, tDel.y, tDel.z};
        float ts[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float fl = floorf(pf[a]);
            ts[a] = ((fl - pf[a] + 0.5f) * float(ps[a]) + 0.5f) * td[a] + t.x;
            p[a] = int(fl);
        }
        tSide = mk3(ts[0], ts[1], ts[2]);
    }
    __device__ void next() {
        m[0] = int((tSide.x < tSide.y) & (tSide.x <= tSide.z));
        m[1] = int((tSide.y < tSide.z) & (tSide.y <= tSide.x));
        m[2] = int((tSide.z < tSide.x) & (tSide.z <= tSide.y));
        t.y = m[0] ? tSide.x : (m[1] ? tSide.y : tSide.z);
    }
    __device__ void step() {
        t.x = t.y;
        tSide = tSide + mk3(float(m[0]) * tDel.x, float(m[1]) * tDel.y, float(m[2]) * tDel.z);
#pragma unroll
        for (int a = 0; a < 3; a++) p[a] += m[a] * ps[a];
    }
};

template <typename AT>
__device__ inline const AT* brick_base(const Tree& t, const AT* atlas, const Node* n) {
    return atlas + slot_of(t, n->mValue) * u64(t.brick_vox);
}

__device__ inline bool inside(f3 p, float r0) {
    return p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0;
}

// rayDeepBrick: emission-absorption march through one brick.  U samples are
// fetched ahead (positions advanced one step at a time, as the reference
// does), then accumulated in order under the reference's loop conditions; a
// sample past the end is fetched but never used.
template <typename AT, int U>
__device__ void deep_brick(const RenderArgs& A, const AT* atlas, int nodeid, f3 tt, f3 pos, f3 dir, float clr[4]) {
    const Tree& t = A.t;
    const Node* n = t.nodes[0] + nodeid;
    const AT* b = brick_base(t, atlas, n);
    const int br = t.brick_res;
    const Corner K = brick_corner<AT>(t, n);
    const uint32_t* cb = A.clr ? brick_base(t, A.clr, n) : nullptr;
    const float dstep = A.s.steps[0];
    f3 vmin = mk3(float(n->mPos[0]), float(n->mPos[1]), float(n->mPos[2]));
    float tx = dstep * ceilf(tt.x / dstep);
    f3 p = pos + dir * tx - vmin;
    f3 wpt = dir * dstep;
    float r0 = float(t.res[0]);
    const float extinct = A.s.extinct[0], albedo = A.s.extinct[1];
    const float minval = A.s.cutoff[0], alphacut = A.s.cutoff[1];
    const float th = A.s.thresh[0], vr = A.s.thresh[2] - A.s.thresh[1];
    for (int it = 0;;) {
        Taps T[U];
        bool ok[U];
        f3 q = p;
#pragma unroll
        for (int u = 0; u < U; u++) {
            ok[u] = it + u < MAX_ITER && inside(q, r0);
            tri_fetch<AT>(b, br, K, q.x + K.o[0], q.y + K.o[1], q.z + K.o[2], T[u]);
            q = q + wpt;
        }
        float raw[U];
        float4 val[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            raw[u] = tri_blend(T[u]);
            float x = fminf(1.0f, fmaxf(0.0f, (raw[u] - th) / vr));
            val[u] = A.tf[int(x * 16300.0f)];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!(clr[3] > alphacut) || !ok[u]) goto done;
            if (raw[u] >= minval) {
                // the reference's products, left to right; the colour channel's
                // factor only when there is one (a factor of 1 changes nothing)
                float w = expf(extinct * val[u].w * dstep);
                float kx = val[u].x * clr[3] * (1.0f - w) * albedo, ky = val[u].y * clr[3] * (1.0f - w) * albedo,
                      kz = val[u].z * clr[3] * (1.0f - w) * albedo;
                if (cb) {
                    const float4 h = color_at(cb, br, K, p.x + K.o[0], p.y + K.o[1], p.z + K.o[2]);
                    kx = kx * h.x;
                    ky = ky * h.y;
                    kz = kz * h.z;
                }
                clr[0] += kx;
                clr[1] += ky;
                clr[2] += kz;
                clr[3] *= w;
            }
            p = p + wpt;
            it++;
        }
    }
done:
    clr[0] = fminf(clr[0], 1.0f); clr[1] = fminf(clr[1], 1.0f); clr[2] = fminf(clr[2], 1.0f);
    clr[3] = fmaxf(clr[3], 0.0f);
}

// raySurfaceTrilinearBrick: first sample >= threshold, gradient normal; U
// samples fetched ahead as in deep_brick.
template <typename AT, int U>
__device__ void surf_brick(const RenderArgs& A, const AT* atlas, int nodeid, f3 tt, f3 pos, f3 dir, f3& hit, f3& nrm,
                           float hclr[4]) {
    const Tree& t = A.t;
    const Node* n = t.nodes[0] + nodeid;
    const AT* b = brick_base(t, atlas, n);
    const Corner K = brick_corner<AT>(t, n);
    const float dstep = A.s.steps[0];
    f3 vmin = mk3(float(n->mPos[0]), float(n->mPos[1]), float(n->mPos[2]));
    float tx = dstep * ceilf(tt.x / dstep);
    f3 p = pos + dir * tx - vmin;
    float r0 = float(t.res[0]);
    const float th = A.s.thresh[0];
    const int br = t.brick_res;
    for (int it = 0;;) {
        Taps T[U];
        bool ok[U];
        f3 q = p;
#pragma unroll
        for (int u = 0; u < U; u++) {
            ok[u] = it + u < MAX_ITER && inside(q, r0);
            tri_fetch<AT>(b, br, K, q.x + K.o[0], q.y + K.o[1], q.z + K.o[2], T[u]);
            q = q + dir * dstep;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (!ok[u]) return;
            if (tri_blend(T[u]) >= th) {
                hit = p + vmin;
                // getGradient at the atlas position p + o
                const f3 P = mk3(p.x + K.o[0], p.y + K.o[1], p.z + K.o[2]);
                f3 g;
                g.x = tri_edge<AT>(b, br, K, P.x - 0.5f, P.y, P.z) - tri_edge<AT>(b, br, K, P.x + 0.5f, P.y, P.z);
                g.y = tri_edge<AT>(b, br, K, P.x, P.y - 0.5f, P.z) - tri_edge<AT>(b, br, K, P.x, P.y + 0.5f, P.z);
                g.z = tri_edge<AT>(b, br, K, P.x, P.y, P.z - 0.5f) - tri_edge<AT>(b, br, K, P.x, P.y, P.z + 0.5f);
                nrm = norm3(g);
                if (A.clr) {
                    const float4 h = color_at(brick_base(t, A.clr, n), br, K, P.x, P.y, P.z);
                    hclr[0] = h.x; hclr[1] = h.y; hclr[2] = h.z; hclr[3] = h.w;
                }
                return;
            }
            p = p + dir * dstep;
            it++;
        }
    }
}

// rayCast.  MODE 0: deep (EA), MODE 1: trilinear surface.
// Per-lane level stacks live in LDS ([level][lane], 64 lanes per workgroup)
// rather than in dynamically indexed registers.
struct LevelStack {
    int nodeid[MAXLEV][64];
    float tMax[MAXLEV][64];
};

template <int MODE, typename AT, int U>
__device__ void ray_cast(const RenderArgs& A, const AT* atlas, f3 pos, f3 dir, f3& hit, f3& nrm, float clr[4],
                         LevelStack& S) {
    const Tree& t = A.t;
    const int lane = threadIdx.x;
    int (&nodeid)[MAXLEV][64] = S.nodeid;
    float (&tMax)[MAXLEV][64] = S.tMax;
    int lev = t.top_lev;
    nodeid[lev][lane] = 0;
    float h0 = (t.bmin[0] - pos.x) / dir.x, h1 = (t.bmax[0] - pos.x) / dir.x;
    float h2 = (t.bmin[1] - pos.y) / dir.y, h3 = (t.bmax[1] - pos.y) / dir.y;
    float h4 = (t.bmin[2] - pos.z) / dir.z, h5 = (t.bmax[2] - pos.z) / dir.z;
    float t0 = fmaxf(fmaxf(fminf(h0, h1), fminf(h2, h3)), fminf(h4, h5));
    float t1 = fminf(fminf(fmaxf(h0, h1), fmaxf(h2, h3)), fmaxf(h4, h5));
    t0 = (t0 < 0) ? 0.0f : t0;
    if (t1 < t0 || t1 < 0) return;
    const Node* node = t.nodes[lev];
    Hdda d;
    d.pos = pos;
    d.dir = dir;
    d.ps[0] = dir.x > 0 ? 1 : -1; d.ps[1] = dir.y > 0 ? 1 : -1; d.ps[2] = dir.z > 0 ? 1 : -1;
    d.t = mk3(t0 + t.epsilon, t1, 0.0f);
    tMax[lev][lane] = t1 - t.epsilon;
    d.prepare(mk3(float(node->mPos[0]), float(node->mPos[1]), float(node->mPos[2])), float(t.range[lev] / t.res[lev]));
    for (int it = 0; it < MAX_ITER && lev > 0 && lev <= t.top_lev && d.p[0] >= 0 && d.p[1] >= 0 && d.p[2] >= 0 &&
                     d.p[0] <= t.res[lev] && d.p[1] <= t.res[lev] && d.p[2] <= t.res[lev]; it++) {
        d.next();
        int res = t.res[lev];
        int child = -1;
        if (d.p[0] < res && d.p[1] < res && d.p[2] < res) {
            int b = (((d.p[2] << t.logdim[lev]) + d.p[1]) << t.logdim[lev]) + d.p[0];
            child = child_of(t, node, b);
        }
        if (child >= 0) {
            if (lev == 1) {
                f3 tt = d.t;
                tt.x += t.epsilon;
                if (MODE == 0) {
                    deep_brick<AT, U>(A, atlas, child, tt, pos, dir, clr);
                    if (clr[3] <= 0) { clr[3] = 0; return; }
                } else {
                    surf_brick<AT, U>(A, atlas, child, tt, pos, dir, hit, nrm, clr);
                    if (hit.z != NOHIT) return;
                }
                d.step();
            } else {
                lev--;
                nodeid[lev][lane] = child;
                node = t.nodes[lev] + child;
                d.t.x += t.epsilon;
                tMax[lev][lane] = d.t.y - t.epsilon;
                d.prepare(mk3(float(node->mPos[0]), float(node->mPos[1]), float(node->mPos[2])),
                          float(t.range[lev] / t.res[lev]));
            }
        } else {
            d.step();
        }
        while (lev <= t.top_lev && d.t.x > tMax[lev][lane]) {
            lev++;
            if (lev <= t.top_lev) {
                node = t.nodes[lev] + nodeid[lev][lane];
                d.prepare(mk3(float(node->mPos[0]), float(node->mPos[1]), float(node->mPos[2])),
                          float(t.range[lev] / t.res[lev]));
            }
        }
    }
}

__device__ inline f3 mmult(const float* m, f3 v) {
    return mk3(v.x * m[0] + v.y * m[4] + v.z * m[8] + m[12], v.x * m[1] + v.y * m[5] + v.z * m[9] + m[13],
               v.x * m[2] + v.y * m[6] + v.z * m[10] + m[14]);
}

template <int MODE, typename AT, int U>
__device__ inline void k_render_body(const RenderArgs& A, uint32_t bid, LevelStack& S) {
    // the workgroup's ray packet: a tile_w x (64 / tile_w) pixel tile
    const int tw = A.tile_w, sh = __builtin_ctz(unsigned(tw));
    const int tiles_x = (A.s.width + tw - 1) >> sh;
    const int tile = int(bid);
    const int x = (tile % tiles_x) * tw + int(threadIdx.x & unsigned(tw - 1));
    const int y = (tile / tiles_x) * (64 >> sh) + int(threadIdx.x >> sh);
    if (x >= A.s.width || y >= A.s.height) return;
    const AT* atlas = reinterpret_cast<const AT*>(A.atlas);
    f3 rpos = mmult(A.s.invxform, mk3(A.s.campos[0], A.s.campos[1], A.s.campos[2]));
    float fx = float(x + 0.5f) / float(A.s.width), fy = float(y + 0.5f) / float(A.s.height);
    f3 v = mk3(fx * A.s.camu[0] + fy * A.s.camv[0] + A.s.cams[0], fx * A.s.camu[1] + fy * A.s.camv[1] + A.s.cams[1],
               fx * A.s.camu[2] + fy * A.s.camv[2] + A.s.cams[2]);
    f3 rdir = norm3(mmult(A.s.invxrot, v));
    float res[4];
    if (MODE == 0) {
        float clr[4] = {0, 0, 0, 1};
        f3 hit = mk3(0, 0, NOHIT), nrm = mk3(0, 0, 0);
        ray_cast<0, AT, U>(A, atlas, rpos, rdir, hit, nrm, clr, S);
        float a = 1.0f - clr[3];
        for (int k = 0; k < 3; k++) res[k] = A.s.backclr[k] + a * (clr[k] - A.s.backclr[k]);
        res[3] = a;
    } else {
        float clr[4] = {1, 1, 1, 1};
        f3 hit = mk3(NOHIT, NOHIT, NOHIT), nrm = mk3(0, 0, 0);
        ray_cast<1, AT, U>(A, atlas, rpos, rdir, hit, nrm, clr, S);
        if (hit.z == NOHIT) {
            for (int k = 0; k < 4; k++) res[k] = A.s.backclr[k];
        } else {
            f3 lp = mk3(A.s.light_pos[0], A.s.light_pos[1], A.s.light_pos[2]);
            f3 ld = norm3(lp - hit);
            float diff = 0.9f * fmaxf(0.0f, dot3(nrm, ld));
            const float amb = 0.1f;
            if (A.s.shadow[0] > 0) {
                f3 h2 = mk3(0, 0, NOHIT), n2 = mk3(0, 0, 0);
                float c2[4] = {0, 0, 0, 1};
                ray_cast<1, AT, U>(A, atlas, hit + nrm * A.s.shadow[1], ld, h2, n2, c2, S);
                diff = (h2.z == NOHIT) ? diff : diff * (1.0f - A.s.shadow[0]);
            }
            for (int k = 0; k < 3; k++) res[k] = clr[k] * (diff + amb);
            res[3] = 1.0f;
        }
    }
    u64 pix = u64(y) * A.s.width + x;
    if (A.out_float) {
        reinterpret_cast<float4*>(A.out)[pix] = make_float4(res[0], res[1], res[2], res[3]);
    } else {
        uchar4 c;
        c.x = (unsigned char)(res[0] * 255.0f);
        c.y = (unsigned char)(res[1] * 255.0f);
        c.z = (unsigned char)(res[2] * 255.0f);
        c.w = (unsigned char)(res[3] * 255.0f);
        reinterpret_cast<uchar4*>(A.out)[pix] = c;
    }
}

// One 64-thread workgroup per pixel tile; several grids' images in one launch.
// Occupancy asked of the compiler per unroll (waves per SIMD: 8, 6, 4 for
// U = 1, 2, 4): the most each variant reaches with no register spills.
constexpr int render_waves(int u) { return u == 1 ? 8 : (u == 2 ? 6 : 4); }
template <int MODE, typename AT, int U>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(render_waves(U)))) void k_render(RenderArgs a0,
                                                                                                     MultiMap m) {
    __shared__ LevelStack S;
    if (m.G <= 1) {
        k_render_body<MODE, AT, U>(a0, blockIdx.x, S);
        return;
    }
    const int g = multi_grid(m, blockIdx.x);
    k_render_body<MODE, AT, U>(multi_rec<RenderArgs>(m, g), blockIdx.x - multi_first(m)[g], S);
}

// Samples fetched ahead per brick step and the ray-packet shape come from
// the launch-shape knobs (tune.cpp): defaults are the measured winners, the
// autotuner re-picks them for the caller's scene.
template <int MODE, typename AT>
static void launch_u(int u, dim3 grid, hipStream_t s, const RenderArgs& a0, const MultiMap& m) {
    if (u == 4) hipLaunchKernelGGL((k_render<MODE, AT, 4>), grid, dim3(64), 0, s, a0, m);
    else if (u == 2) hipLaunchKernelGGL((k_render<MODE, AT, 2>), grid, dim3(64), 0, s, a0, m);
    else hipLaunchKernelGGL((k_render<MODE, AT, 1>), grid, dim3(64), 0, s, a0, m);
}

hipError_t launch_render_multi(const RenderArgs* a, int G, Arena* ar, hipStream_t s) {
    for (int variant = 0; variant < 4; variant++) {
        const int mode = variant & 1;
        const bool half = variant >> 1;
        std::vector<RenderArgs> sel;
        std::vector<uint32_t> nb;
        for (int g = 0; g < G; g++) {
            if (a[g].shading != SHADE_VOLUME && a[g].shading != SHADE_TRILINEAR) return hipErrorInvalidValue;
            const int am = a[g].shading == SHADE_VOLUME ? 0 : 1;
            if (am != mode || (a[g].atlas_type == T_HALF) != half) continue;
            const int tw = knob(mode ? KNOB_RENDER_TILE_ISO : KNOB_RENDER_TILE_EA), th = 64 / tw;
            sel.push_back(a[g]);
            sel.back().tile_w = tw;
            nb.push_back(uint32_t(((a[g].s.width + tw - 1) / tw) * ((a[g].s.height + th - 1) / th)));
        }
        if (sel.empty()) continue;
        MultiLaunch L;
        hipError_t e = multi_setup(ar, sel.data(), nb.data(), int(sel.size()), s, L);
        if (e != hipSuccess) return e;
        if (L.blocks == 0) continue;
        const int u = knob(mode ? KNOB_RENDER_UNROLL_ISO : KNOB_RENDER_UNROLL_EA);
        if (mode == 0 && half) launch_u<0, __half>(u, dim3(L.blocks), s, sel[0], L.map);
        else if (mode == 0) launch_u<0, float>(u, dim3(L.blocks), s, sel[0], L.map);
        else if (half) launch_u<1, __half>(u, dim3(L.blocks), s, sel[0], L.map);
        else launch_u<1, float>(u, dim3(L.blocks), s, sel[0], L.map);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gvdb
