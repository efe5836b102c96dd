// camera.cpp -- pinhole camera to the render scene block.
//
// Camera3D::setOrbit places the eye at to + dist * (cos(ay) sin(ax),
// sin(ay), cos(ay) cos(ax)) (src/gvdb_camera.cpp:90-105 of the reference);
// updateMatricies builds a gluLookAt basis and a gluPerspective projection
// (:175-228); PrepareRender (src/gvdb_volume_gvdb.cpp:4254-4306) sends the
// eye and the near-plane corner rays of inverse(proj * view-rotation)
// (updateFrustum / inverseRayProj, src/gvdb_camera.cpp:292-346, :382-393).
// The aspect ratio is the render buffer's (Scene::SetRes).
#include "volume.h"
#include "../../include/gvdb_amd.h"
#include <cmath>
#include <cstring>

namespace gvdb {

namespace {

void cross_f(const float* a, const float* b, float* o) {
    double ax = a[0], ay = a[1], az = a[2];
    o[0] = float(ay * double(b[2]) - az * double(b[1]));
    o[1] = float(-ax * double(b[2]) + az * double(b[0]));
    o[2] = float(ax * double(b[1]) - ay * double(b[0]));
}

void normalize_f(float* v) {
    double n = double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2];
    if (n != 0.0) {
        double r = 1.0 / std::sqrt(n);
        for (int a = 0; a < 3; a++) v[a] = float(v[a] * r);
    }
}

// column-major 4x4: r = a * b
void mat_mul(const float* a, const float* b, float* r) {
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            float s = 0;
            for (int k = 0; k < 4; k++) s += a[k * 4 + row] * b[c * 4 + k];
            r[c * 4 + row] = s;
        }
}

// General 4x4 inverse by Gauss-Jordan elimination with partial pivoting in double.
bool mat_inverse(const float* m, float* out) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
        double inv = 1.0 / a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            double f = a[r][c];
            if (f != 0.0)
                for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[c * 4 + r] = float(a[r][c + 4]);
    return true;
}

}  // namespace

void Volume::scene_block(int w, int h, ScnBlock& s) const {
    memset(&s, 0, sizeof(s));
    s.width = w;
    s.height = h;
    const float d2r = 3.141592f / 180.0f;
    double dx = std::cos(double(cam.angs[1] * d2r)) * std::sin(double(cam.angs[0] * d2r));
    double dy = std::sin(double(cam.angs[1] * d2r));
    double dz = std::cos(double(cam.angs[1] * d2r)) * std::cos(double(cam.angs[0] * d2r));
    float from[3] = {cam.to[0] + float(dx) * cam.dist, cam.to[1] + float(dy) * cam.dist, cam.to[2] + float(dz) * cam.dist};
    float dir[3] = {cam.to[0] - from[0], cam.to[1] - from[1], cam.to[2] - from[2]};
    normalize_f(dir);
    const float up[3] = {0, 1, 0};
    float side[3], upv[3];
    cross_f(dir, up, side);
    normalize_f(side);
    cross_f(side, dir, upv);
    normalize_f(upv);
    // view rotation (basis rows side, up, -dir), translation dropped
    float view[16] = {side[0], upv[0], -dir[0], 0, side[1], upv[1], -dir[1], 0, side[2], upv[2], -dir[2], 0, 0, 0, 0, 1};
    float aspect = float(w) / float(h);
    float sx = float(std::tan(cam.fov * d2r / 2.0f)) * cam.znear;
    float sy = sx / aspect;
    float proj[16] = {0};
    proj[0] = 2.0f * cam.znear / sx;
    proj[5] = 2.0f * cam.znear / sy;
    proj[10] = -(cam.zfar + cam.znear) / (cam.zfar - cam.znear);
    proj[14] = -(2.0f * cam.zfar * cam.znear) / (cam.zfar - cam.znear);
    proj[11] = -1.0f;
    float pv[16], ipv[16];
    mat_mul(proj, view, pv);
    if (!mat_inverse(pv, ipv)) memset(ipv, 0, sizeof(ipv));
    auto ray = [&](float x, float y, float z, float* o) {
        float q[4];
        for (int r = 0; r < 4; r++) q[r] = ipv[r] * x + ipv[4 + r] * y + ipv[8 + r] * z + ipv[12 + r];
        o[0] = q[0] / q[3]; o[1] = q[1] / q[3]; o[2] = q[2] / q[3];
    };
    float tl[3], tr[3], bl[3];
    ray(-1.0f, 1.0f, cam.znear, tl);
    ray(1.0f, 1.0f, cam.znear, tr);
    ray(-1.0f, -1.0f, cam.znear, bl);
    for (int a = 0; a < 3; a++) {
        s.campos[a] = from[a];
        s.cams[a] = tl[a];
        s.camu[a] = tr[a] - tl[a];
        s.camv[a] = bl[a] - tl[a];
        s.light_pos[a] = light_[a];
        s.extinct[a] = extinct_[a];
        s.steps[a] = steps_[a];
        s.cutoff[a] = cutoff_[a];
        s.thresh[a] = vrange_[a];
        s.shadow[a] = shadow_[a];
    }
    for (int k = 0; k < 4; k++) s.backclr[k] = back_[k];
    for (int k = 0; k < 16; k++) {
        s.invxform[k] = xf_inv_[k];
        s.invxrot[k] = xf_invrot_[k];
    }
    // the light from application space into index space, as a position
    // (PrepareRender, gvdb_volume_gvdb.cpp:4271-4275)
    const float* m = xf_inv_;
    for (int a = 0; a < 3; a++)
        s.light_pos[a] = light_[0] * m[a] + light_[1] * m[4 + a] + light_[2] * m[8 + a] + m[12 + a];
}

// SetTransform: xform = T R S PT with R = Rz Ry Rx (angles in degrees about
// x, then y, then z, right-handed), S = diag(scale), PT a translation by
// pretrans applied first.  Its inverse PT^-1 S^-1 R^T T^-1 and the inverse of
// the linear part, S^-1 R^T, are formed in double and rounded to float (the
// reference inverts the float matrix with a general 4x4 inverse in double).
int Volume::set_transform(const float pretrans[3], const float scale[3], const float angs[3], const float trans[3]) {
    for (int a = 0; a < 3; a++)
        if (!(scale[a] != 0.0f) || !std::isfinite(scale[a]) || !std::isfinite(angs[a]) || !std::isfinite(trans[a]) ||
            !std::isfinite(pretrans[a]))
            return fail(GVDB_ERR_INVALID, "SetTransform: finite values and nonzero scales");
    for (int a = 0; a < 3; a++) {
        xf_pretrans_[a] = pretrans[a];
        xf_scale_[a] = scale[a];
        xf_angs_[a] = angs[a];
        xf_trans_[a] = trans[a];
    }
    const double d2r = 3.14159265358979323846 / 180.0;
    const double cx = std::cos(angs[0] * d2r), sx = std::sin(angs[0] * d2r), cy = std::cos(angs[1] * d2r),
                 sy = std::sin(angs[1] * d2r), cz = std::cos(angs[2] * d2r), sz = std::sin(angs[2] * d2r);
    // R = Rz Ry Rx, row-major here: R[r][c]
    const double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
    const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    const double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
    double RyRx[3][3], R[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            RyRx[r][c] = 0;
            for (int k = 0; k < 3; k++) RyRx[r][c] += Ry[r][k] * Rx[k][c];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            R[r][c] = 0;
            for (int k = 0; k < 3; k++) R[r][c] += Rz[r][k] * RyRx[k][c];
        }
    // L = S^-1 R^T (the inverse of R S), then inv = L (x - t) - pt
    double L[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) L[r][c] = R[c][r] / double(scale[r]);
    double lt[3];
    for (int r = 0; r < 3; r++) lt[r] = L[r][0] * trans[0] + L[r][1] * trans[1] + L[r][2] * trans[2];
    for (int k = 0; k < 16; k++) xf_inv_[k] = xf_invrot_[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) xf_inv_[c * 4 + r] = xf_invrot_[c * 4 + r] = float(L[r][c]);
        xf_inv_[12 + r] = float(-lt[r] - double(pretrans[r]));
    }
    return GVDB_OK;
}

}  // namespace gvdb
