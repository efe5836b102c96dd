// tune.cpp -- launch-shape knobs and the autotuner that picks them.
//
// The knobs are the per-kernel launch shapes the path leaves open: samples
// fetched ahead and ray-packet shape of the march, threads per brick of the
// gather, bricks per workgroup of the multi-grid apron.  None changes a
// result (every value computes the same bits); they only move time.  The
// tuner runs the caller's own workload once per value (a warm-up, then the
// median of three timed runs by HIP events on the grid's or batch's stream)
// and keeps the fastest.  Values are process-wide: tuning one grid sets the
// shape every later launch uses.
#include "kernels.h"
#include "volume.h"
#include "batch.h"
#include "../../include/gvdb_amd.h"
#include <algorithm>
#include <cfloat>

namespace gvdb {

static const int kValues[KNOB_COUNT][KNOB_MAX_VALUES] = {
    {1, 2, 4}, {1, 2, 4}, {8, 16, 32, 64}, {8, 16, 32, 64}, {128, 256}, {1, 2, 4, 8}, {0, 1}, {0, 1}, {0, 1, 2, 3, 4}, {1, 2, 4, 8, 16, 32}};
static const int kCount[KNOB_COUNT] = {3, 3, 4, 4, 2, 4, 2, 2, 5, 6};
// defaults: the values measured fastest on the medium and 64-grid workloads
static int g_knob[KNOB_COUNT] = {4, 2, 8, 8, 128, 4, 0, 1, 0, 8};

int knob(int k) { return (k >= 0 && k < KNOB_COUNT) ? g_knob[k] : 0; }

int knob_values(int k, int out[KNOB_MAX_VALUES]) {
    int n = 0;
    if (k < 0 || k >= KNOB_COUNT) return 0;
    for (int i = 0; i < kCount[k]; i++) out[n++] = kValues[k][i];
    return n;
}

bool set_knob(int k, int v) {
    int vals[KNOB_MAX_VALUES];
    const int n = knob_values(k, vals);
    for (int i = 0; i < n; i++)
        if (vals[i] == v) {
            g_knob[k] = v;
            return true;
        }
    return false;
}

// Median of three timed runs after one warm-up; run(ms) returns a status.
template <class F>
static int time_runs(F&& run, float& out) {
    float t[3];
    for (int rep = 0; rep < 4; rep++) {
        float ms = 0.f;
        int rc = run(ms);
        if (rc != GVDB_OK) return rc;
        if (rep) t[rep - 1] = ms;
    }
    std::sort(t, t + 3);
    out = t[1];
    return GVDB_OK;
}

// Tries every value of the knobs in `ks` (all combinations), keeps the
// fastest; ms[] gets one time per combination, the first knob varying
// slowest.  On failure the knobs are restored.
template <class F>
static int tune_knobs(const int* ks, int nk, F&& run, int* best, float* ms) {
    int vals[2][KNOB_MAX_VALUES], cnt[2] = {1, 1}, keep[2] = {0, 0};
    for (int i = 0; i < nk; i++) {
        cnt[i] = knob_values(ks[i], vals[i]);
        keep[i] = g_knob[ks[i]];
    }
    float best_ms = FLT_MAX;
    int pick[2] = {keep[0], keep[1]};
    int idx = 0;
    for (int a = 0; a < cnt[0]; a++)
        for (int b = 0; b < (nk > 1 ? cnt[1] : 1); b++, idx++) {
            g_knob[ks[0]] = vals[0][a];
            if (nk > 1) g_knob[ks[1]] = vals[1][b];
            float t = 0.f;
            int rc = time_runs(run, t);
            if (rc != GVDB_OK) {
                for (int i = 0; i < nk; i++) g_knob[ks[i]] = keep[i];
                return rc;
            }
            if (ms) ms[idx] = t;
            if (t < best_ms) {
                best_ms = t;
                pick[0] = vals[0][a];
                if (nk > 1) pick[1] = vals[1][b];
            }
        }
    for (int i = 0; i < nk; i++) {
        g_knob[ks[i]] = pick[i];
        if (best) best[i] = pick[i];
    }
    return GVDB_OK;
}

int tune_render(Volume& v, int shading, int chan, int buf, int* best2, float* ms12) {
    if (shading != GVDB_SHADE_VOLUME && shading != GVDB_SHADE_TRILINEAR) return GVDB_ERR_INVALID;
    const bool ea = shading == GVDB_SHADE_VOLUME;
    const int ks[2] = {ea ? KNOB_RENDER_TILE_EA : KNOB_RENDER_TILE_ISO, ea ? KNOB_RENDER_UNROLL_EA : KNOB_RENDER_UNROLL_ISO};
    int b[2];
    int rc = tune_knobs(ks, 2, [&](float& ms) {
        int r = v.render(shading, chan, buf);
        float k[4];
        if (r == GVDB_OK) r = v.kernel_ms(k);
        ms = k[2];
        return r;
    }, b, ms12);
    if (rc == GVDB_OK && best2) {
        best2[0] = b[1];   // unroll
        best2[1] = b[0];   // tile width
    }
    return rc;
}

int tune_render_unroll(Volume& v, int shading, int chan, int buf, int* best, float* ms3) {
    if (shading != GVDB_SHADE_VOLUME && shading != GVDB_SHADE_TRILINEAR) return GVDB_ERR_INVALID;
    const int ks[1] = {shading == GVDB_SHADE_VOLUME ? KNOB_RENDER_UNROLL_EA : KNOB_RENDER_UNROLL_ISO};
    return tune_knobs(ks, 1, [&](float& ms) {
        int r = v.render(shading, chan, buf);
        float k[4];
        if (r == GVDB_OK) r = v.kernel_ms(k);
        ms = k[2];
        return r;
    }, best, ms3);
}

int tune_gather(Volume& v, int scs, int n, float radius, const float trans[3], int len, int chan, int* best,
                float* ms2) {
    const int ks[1] = {KNOB_GATHER_THREADS};
    return tune_knobs(ks, 1, [&](float& ms) {
        int r = v.gather_density(scs, n, radius, trans, len, chan, 0, false);
        float k[4];
        if (r == GVDB_OK) r = v.kernel_ms(k);
        ms = k[0];
        return r;
    }, best, ms2);
}

// Batched stages: 0 = GatherDensity (threads per brick), 1 = UpdateApron
// (bricks per workgroup), 2 = Render (ray-packet shape x samples ahead).
int tune_batch(Batch& B, Volume* const* vs, int G, int stage, int scs, const int* len, const float* radius, int chan,
               float boundval, const int* shading, const int* chans, const int* bufs, int* best, float* ms,
               int* failed) {
    auto timed = [&](auto&& call) {
        return [&, call](float& ms) {
            int r = B.synchronize();
            if (r == GVDB_OK) r = call();
            if (r == GVDB_OK) r = B.elapsed_ms(&ms);
            return r;
        };
    };
    if (stage == 0) {
        const int ks[1] = {KNOB_GATHER_THREADS};
        return tune_knobs(ks, 1, timed([&] { return B.gather_density(vs, G, scs, len, radius, chan, false, failed); }),
                          best, ms);
    }
    if (stage == 1) {
        // static runs of 1, 2, 4, 8 bricks, then the persistent queue
        const int ks[1] = {KNOB_APRON_RUN};
        const int keepq = g_knob[KNOB_APRON_QUEUE];
        g_knob[KNOB_APRON_QUEUE] = 0;
        int b0 = 0;
        int rc = tune_knobs(ks, 1, timed([&] { return B.update_apron(vs, G, chan, boundval, failed); }), &b0, ms);
        if (rc != GVDB_OK) {
            g_knob[KNOB_APRON_QUEUE] = keepq;
            return rc;
        }
        g_knob[KNOB_APRON_QUEUE] = 1;
        float tq = 0.f;
        rc = time_runs(timed([&] { return B.update_apron(vs, G, chan, boundval, failed); }), tq);
        if (rc != GVDB_OK) {
            g_knob[KNOB_APRON_QUEUE] = keepq;
            return rc;
        }
        if (ms) ms[4] = tq;
        int vals[KNOB_MAX_VALUES];
        const int n = knob_values(KNOB_APRON_RUN, vals);
        float tb = FLT_MAX;
        for (int i = 0; i < n && ms; i++)
            if (vals[i] == b0) tb = ms[i];
        g_knob[KNOB_APRON_QUEUE] = (ms && tq < tb) ? 1 : 0;
        if (best) {
            best[0] = b0;
            best[1] = g_knob[KNOB_APRON_QUEUE];
        }
        return GVDB_OK;
    }
    if (stage == 2) {
        if (G <= 0 || (shading[0] != GVDB_SHADE_VOLUME && shading[0] != GVDB_SHADE_TRILINEAR)) return GVDB_ERR_INVALID;
        const bool ea = shading[0] == GVDB_SHADE_VOLUME;
        const int ks[2] = {ea ? KNOB_RENDER_TILE_EA : KNOB_RENDER_TILE_ISO,
                           ea ? KNOB_RENDER_UNROLL_EA : KNOB_RENDER_UNROLL_ISO};
        int b[2];
        int rc = tune_knobs(ks, 2, timed([&] { return B.render(vs, G, shading, chans, bufs, failed); }), b, ms);
        if (rc == GVDB_OK && best) {
            best[0] = b[1];
            best[1] = b[0];
        }
        return rc;
    }
    return GVDB_ERR_INVALID;
}

}  // namespace gvdb
