// serve.cpp -- see serve.h.
#include "serve.h"
#include "../../include/gvdb_amd.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <thread>

namespace gvdb {

static const int STAGE_BYTES = 1024 * 1024 * 4;   // largest render (1024^2 RGBA8)
// device bytes an insert of one point can add to its grid: the point, its
// bucket key/slot, its bucketed copy and about a dozen 12-byte subcell entries
static const unsigned long long INSERT_BYTES_PER_POINT = 12 + 8 + 12 + 12 * 14;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Server::Server(int device, unsigned long long budget, int max_batch)
    : device_(device), budget_(budget), max_batch_(max_batch) {}

Server::~Server() {
    (void)hipSetDevice(device_);
    if (ib_) (void)ib_->synchronize();
    if (rb_) (void)rb_->synchronize();
    for (auto p : stage_) (void)hipHostFree(p);
    if (rev_) (void)hipEventDestroy(rev_);
    for (auto& g : grids_) g.pts.release();
}

int Server::init() {
    if (max_batch_ < 1 || max_batch_ > 1024) { err = "server: max_batch must be 1..1024"; return GVDB_ERR_INVALID; }
    (void)hipSetDevice(device_);
    ib_.reset(new Batch(device_, 1024));
    rb_.reset(new Batch(device_, 1024));
    int rc = ib_->init();
    if (!rc) rc = rb_->init();
    if (rc) { err = "server: " + ib_->err + rb_->err; return rc; }
    stage_.resize(max_batch_, nullptr);
    for (auto& p : stage_) {
        if (hipHostMalloc(reinterpret_cast<void**>(&p), STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            err = "server: pinned readback slots";
            return GVDB_ERR_OOM;
        }
    }
    if (hipEventCreateWithFlags(&rev_, hipEventDisableTiming) != hipSuccess) { err = "server: event"; return GVDB_ERR_HIP; }
    return GVDB_OK;
}

unsigned long long Server::device_bytes() const {
    unsigned long long b = 0;
    for (const auto& g : grids_) b += g.v->device_bytes() + g.pts.bytes;
    return b;
}

int Server::append(Grid& g, const float* host, int n, hipStream_t s) {
    if (n <= 0) return GVDB_OK;
    if (g.n > 0x7FFFFFFF - n) { err = "server: grid point count overflows"; return GVDB_ERR_RANGE; }
    size_t need = size_t(g.n + n) * 12;
    hipError_t e = g.pts.ensure(need + need / 2, true, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(reinterpret_cast<char*>(g.pts.p) + size_t(g.n) * 12, host, size_t(n) * 12,
                           hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        err = std::string("server: points upload: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? GVDB_ERR_OOM : GVDB_ERR_HIP;
    }
    g.n += n;
    return g.v->set_points_gpu(g.pts.p, g.n, 0, 12);
}

int Server::add_grid(Volume* v, const float* host, int n, float radius) {
    if (!v || !host || n <= 0) { err = "server: grid needs a volume and points"; return GVDB_ERR_INVALID; }
    if (v->device != device_) { err = "server: grid on another device"; return GVDB_ERR_INVALID; }
    Grid g;
    g.v = v;
    g.radius = radius;
    g.init_host.assign(host, host + size_t(n) * 3);
    g.n0 = n;
    float lo[3] = {host[0], host[1], host[2]}, hi[3] = {host[0], host[1], host[2]};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], host[3 * i + a]);
            hi[a] = std::max(hi[a], host[3 * i + a]);
        }
    for (int a = 0; a < 3; a++) {
        g.centre[a] = 0.5f * (lo[a] + hi[a]);
        g.extent = std::max(g.extent, hi[a] - lo[a]);
    }
    // render buffer sized once for the largest request
    int rc = v->add_render_buf(0, 1024, 1024, 4);
    if (rc) { err = v->err; return rc; }
    grids_.push_back(std::move(g));
    return int(grids_.size()) - 1;
}

int Server::build(const std::vector<int>& which, Batch& b) {
    if (which.empty()) return GVDB_OK;
    std::vector<Volume*> vs;
    std::vector<int> n, len(which.size());
    std::vector<float> r;
    for (int k : which) {
        vs.push_back(grids_[k].v);
        n.push_back(grids_[k].n);
        r.push_back(grids_[k].radius);
    }
    const int G = int(vs.size());
    std::vector<float> z(size_t(G) * 3, 0.f);
    int f = -1;
    for (Volume* v : vs) v->request_full_rebuild(true);
    int rc = b.rebuild_topology(vs.data(), G, n.data(), r.data(), z.data(), &f);
    if (!rc) rc = b.finish_topology(vs.data(), G, false, true, &f);
    if (!rc) rc = b.update_atlas(vs.data(), G, &f);
    if (!rc) rc = b.clear_channel(vs.data(), G, 0, &f);
    if (!rc) rc = b.insert_points_subcell(vs.data(), G, 4, n.data(), r.data(), z.data(), len.data(), &f);
    if (!rc) rc = b.gather_density(vs.data(), G, 4, len.data(), r.data(), 0, false, &f);
    if (!rc) rc = b.update_apron(vs.data(), G, 0, 0.0f, &f);
    if (!rc) rc = b.synchronize();
    if (rc) err = "server build: " + b.err;
    return rc;
}

int Server::reset() {
    std::vector<int> all(grids_.size());
    std::iota(all.begin(), all.end(), 0);
    for (auto& g : grids_) {
        if (hipStreamSynchronize(g.v->stream) != hipSuccess) { err = "server: stream"; return GVDB_ERR_HIP; }
        int rc = g.v->use_stream(ib_->shared);
        if (rc) { err = g.v->err; return rc; }
    }
    for (auto& g : grids_) {
        g.n = 0;
        int rc = append(g, g.init_host.data(), g.n0, ib_->stream);
        if (rc) return rc;
    }
    for (auto& r : reqs_) { r.done = -1; r.status = 0; r.slot = -1; }
    return build(all, *ib_);
}

int Server::submit_render(int grid, double arrival, int w, int h, int shading, const float angs[3], float dist,
                          unsigned char* out) {
    if (grid < 0 || grid >= int(grids_.size()) || w <= 0 || h <= 0 || w > 1024 || h > 1024 || !out || !angs ||
        (shading != SHADE_VOLUME && shading != SHADE_TRILINEAR)) {
        err = "server: render request needs a grid, 1..1024 pixels per side, a shading mode and an output";
        return GVDB_ERR_INVALID;
    }
    Req r;
    r.kind = 0;
    r.grid = grid;
    r.arrival = arrival;
    r.w = w;
    r.h = h;
    r.shading = shading;
    for (int a = 0; a < 3; a++) r.angs[a] = angs[a];
    r.dist = dist;
    r.out = out;
    reqs_.push_back(r);
    return int(reqs_.size()) - 1;
}

int Server::submit_insert(int grid, double arrival, const float* host, int n) {
    if (grid < 0 || grid >= int(grids_.size()) || !host || n <= 0) {
        err = "server: insert request needs a grid and points";
        return GVDB_ERR_INVALID;
    }
    Req r;
    r.kind = 1;
    r.grid = grid;
    r.arrival = arrival;
    r.pts = host;
    r.n = n;
    reqs_.push_back(r);
    return int(reqs_.size()) - 1;
}

void Server::clear_requests() { reqs_.clear(); }

int Server::render_setup(Req& r) {
    Grid& g = grids_[r.grid];
    Volume* v = g.v;
    for (int a = 0; a < 3; a++) {
        v->cam.angs[a] = r.angs[a];
        v->cam.to[a] = g.centre[a];
    }
    v->cam.dist = r.dist > 0 ? r.dist : 1.5f * g.extent;
    int rc = v->add_render_buf(0, r.w, r.h, 4);   // no reallocation: sized at add_grid
    if (rc) err = v->err;
    return rc;
}

// --------------------------------------------------------------- plain ----
int Server::run_plain(double t0) {
    std::vector<int> order(reqs_.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs_[a].arrival < reqs_[b].arrival; });
    for (int i : order) {
        Req& r = reqs_[i];
        double wait = t0 + r.arrival - now_s();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        Grid& g = grids_[r.grid];
        Volume* v = g.v;
        int rc = v->use_stream(std::shared_ptr<SharedStream>());
        if (r.kind == 0) {
            if (!rc) rc = render_setup(r);
            if (!rc) rc = v->render(r.shading, 0, 0);
            if (!rc) rc = v->read_render_buf(0, r.out);
        } else {
            if (!rc) rc = append(g, r.pts, r.n, v->stream);
            int len = 0;
            const float z[3] = {0, 0, 0};
            if (!rc) rc = v->request_full_rebuild(true);
            if (!rc) rc = v->rebuild_topology(g.n, g.radius, z);
            if (!rc) rc = v->finish_topology(false, true);
            if (!rc) rc = v->update_atlas();
            if (!rc) rc = v->clear_channel(0);
            if (!rc) rc = v->insert_points_subcell(4, g.n, g.radius, z, &len);
            if (!rc) rc = v->gather_density(4, g.n, g.radius, z, len, 0, 0, false);
            if (!rc) rc = v->update_apron(0, 0.0f);
            if (!rc && hipStreamSynchronize(v->stream) != hipSuccess) rc = GVDB_ERR_HIP;
        }
        r.status = rc;
        r.done = now_s() - t0;
    }
    return GVDB_OK;
}

// ---------------------------------------------------------- continuous ----
int Server::run_continuous(double t0) {
    const int N = int(reqs_.size());
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs_[a].arrival < reqs_[b].arrival; });
    std::vector<char> busy(grids_.size(), 0), blocked(grids_.size(), 0);
    std::vector<char> taken(N, 0);
    std::vector<int> inflight_r;                 // renders on the GPU
    int completed = 0;
    size_t head = 0;                             // first not-yet-taken request in arrival order
    while (completed < N) {
        double t = now_s() - t0;
        // renders finished?
        hipError_t q = inflight_r.empty() ? hipErrorNotReady : hipEventQuery(rev_);
        if (q != hipSuccess && q != hipErrorNotReady) {
            err = std::string("server: render batch: ") + hipGetErrorString(q);
            return GVDB_ERR_HIP;
        }
        if (q == hipSuccess) {
            double done = now_s() - t0;
            for (int i : inflight_r) {
                Req& r = reqs_[i];
                memcpy(r.out, stage_[r.slot], size_t(r.w) * r.h * 4);
                r.done = done;
                busy[r.grid] = 0;
                completed++;
            }
            inflight_r.clear();
        }
        // admission, in arrival order; a grid's requests stay in order
        std::fill(blocked.begin(), blocked.end(), 0);
        std::vector<int> renders, inserts;
        unsigned long long extra = 0, base = device_bytes();
        while (head < order.size() && taken[order[head]]) head++;
        for (size_t k = head; k < order.size(); k++) {
            int i = order[k];
            if (taken[i]) continue;
            Req& r = reqs_[i];
            if (r.arrival > t) break;
            if (busy[r.grid] || blocked[r.grid]) { blocked[r.grid] = 1; continue; }
            if (r.kind == 0) {
                if (!inflight_r.empty() || int(renders.size()) >= max_batch_) { blocked[r.grid] = 1; continue; }
                r.slot = int(renders.size());
                renders.push_back(i);
            } else {
                unsigned long long need = INSERT_BYTES_PER_POINT * (unsigned long long)(grids_[r.grid].n + r.n);
                if (base + need > budget_) {
                    // cannot fit even alone: refused
                    r.status = GVDB_ERR_OOM;
                    r.done = now_s() - t0;
                    taken[i] = 1;
                    completed++;
                    continue;
                }
                if (int(inserts.size()) >= max_batch_ || base + extra + need > budget_) {
                    blocked[r.grid] = 1;
                    continue;
                }
                extra += need;
                inserts.push_back(i);
            }
            taken[i] = 1;
            busy[r.grid] = 1;
        }
        if (!renders.empty()) {
            std::vector<Volume*> vs;
            std::vector<int> sh, ch, bf;
            int rc = GVDB_OK;
            for (int i : renders) {
                Req& r = reqs_[i];
                rc = render_setup(r);
                if (rc) break;
                vs.push_back(grids_[r.grid].v);
                sh.push_back(r.shading);
                ch.push_back(0);
                bf.push_back(0);
            }
            int f = -1;
            if (!rc) rc = rb_->render(vs.data(), int(vs.size()), sh.data(), ch.data(), bf.data(), &f);
            for (int i : renders) {
                if (rc) break;
                Req& r = reqs_[i];
                if (hipMemcpyAsync(stage_[r.slot], grids_[r.grid].v->render_buf_ptr(0), size_t(r.w) * r.h * 4,
                                   hipMemcpyDeviceToHost, rb_->stream) != hipSuccess)
                    rc = GVDB_ERR_HIP;
            }
            if (!rc && hipEventRecord(rev_, rb_->stream) != hipSuccess) rc = GVDB_ERR_HIP;
            if (rc) {
                err = "server render: " + rb_->err;
                return rc;
            }
            inflight_r = renders;
        }
        if (!inserts.empty()) {
            std::vector<int> which;
            int rc = GVDB_OK;
            for (int i : inserts) {
                Req& r = reqs_[i];
                Grid& g = grids_[r.grid];
                rc = g.v->use_stream(ib_->shared);
                if (!rc) rc = append(g, r.pts, r.n, ib_->stream);
                if (rc) break;
                which.push_back(r.grid);
            }
            if (!rc) rc = build(which, *ib_);
            double done = now_s() - t0;
            for (int i : inserts) {
                Req& r = reqs_[i];
                r.status = rc;
                r.done = done;
                busy[r.grid] = 0;
                completed++;
            }
            if (rc) return rc;
        }
        if (renders.empty() && inserts.empty()) {
            // idle: poll the renders in flight or wait for the next arrival
            double next = head < order.size() ? reqs_[order[head]].arrival : t;
            double wait = inflight_r.empty() ? std::max(0.0, next - (now_s() - t0)) : 0.0;
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.002)));
            else if (!inflight_r.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
    return GVDB_OK;
}

int Server::run(int mode, double* latency, int* status, double* wall) {
    if (reqs_.empty()) { if (wall) *wall = 0; return GVDB_OK; }
    (void)hipSetDevice(device_);
    for (auto& r : reqs_) { r.done = -1; r.status = 0; r.slot = -1; }
    // all earlier work drains before the clock starts
    for (auto& g : grids_)
        if (hipStreamSynchronize(g.v->stream) != hipSuccess) { err = "server: stream"; return GVDB_ERR_HIP; }
    double t0 = now_s();
    int rc = mode == 0 ? run_plain(t0) : run_continuous(t0);
    if (rc) return rc;
    double first = reqs_[0].arrival, last = 0;
    for (size_t i = 0; i < reqs_.size(); i++) {
        first = std::min(first, reqs_[i].arrival);
        last = std::max(last, reqs_[i].done);
        if (latency) latency[i] = reqs_[i].done - reqs_[i].arrival;
        if (status) status[i] = reqs_[i].status;
    }
    if (wall) *wall = last - first;
    return GVDB_OK;
}

}  // namespace gvdb
