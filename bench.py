"""Benchmark of the point-to-sparse-volume path on one MI355X.

Workload (plain method, medium, on GPU): 4,194,304 points uniform in the
400x200x400-voxel slab [56,456]x[20,220]x[56,456] plus 200,000 points uniform
on the plane y = 10; tree log2 dims {3,3,3,4,5} (8^3 bricks, apron 1), one
FP32 density channel, gather radius 2, subcell 4; isosurface march at density
0.5 with trilinear sampling, 1024x1024, step 0.25, Phong + shadow ray.

One step = RebuildTopology + FinishTopology + UpdateAtlas + ClearChannel +
InsertPointsSubcell + GatherDensity + UpdateApron (the "splat" part) and one
Render.  Inputs are resident in HBM before timing starts.

Prints one JSON line.  `value` = points splatted per second (points / splat
time); `mrays_per_s` = primary rays per second of the render.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

T_START = time.time()

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "points splatted/s + Mrays/s on 1 MI355X; batched p50/p99 latency vs plain form"


def algorithmic_bytes(v, n_points, entries):
    """Per-launch algorithmic HBM bytes of the gather and apron kernels, as
    the path's measurement plan states them (E subcell entries, S = 8 B
    subcells, B bricks, N points):
      gather: E*12 (entries read) + S*8 (subcell counts/offsets) + B*512*4
              (interior writes);
      apron:  B*488*4 read + B*488*4 written.
    `gather_design` is this build's own minimum for its gather, which reads
    each bucketed point once instead of one entry per (point, subcell):
    N*16 + B*(64 + 108 + 2048)."""
    B = len(v.pool(0, 0))
    S = 8 * B
    return dict(B=B, S=S, E=int(entries), gather=int(entries) * 12 + S * 8 + B * 512 * 4, apron=B * 488 * 8,
                gather_design=n_points * 16 + B * (64 + 108 + 2048))


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE passes (profiles/traffic_medium_slab.json, scaled
    by the copy calibration recorded there); None when not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_medium_slab.json")) as f:
            k = json.load(f)["kernels"].get("gvdb::" + kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else float(k["traffic_bytes"])


def run_gpu(args):
    import gvdb_amd as G
    from gvdb_amd import workloads as W
    pts = W.slab_plane()
    n = len(pts)
    sc = W.iso_scene(args.width, args.width)
    v = G.VolumeGVDB(0)
    v.Configure(5, 4, 3, 3, 3)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, G.T_FLOAT, 1)
    v.SetPoints(pts)
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 4)
    r = 2.0
    z = (0.0, 0.0, 0.0)

    def splat():
        v.RequestFullRebuild(True)
        v.RebuildTopology(n, r, z)
        v.FinishTopology(False, True)
        v.UpdateAtlas()
        v.ClearChannel(0)
        ln = v.InsertPointsSubcell(4, n, r, z)
        v.GatherDensity(4, n, r, z, ln, 0, 0, False)
        v.UpdateApron(0, 0.0)
        return ln

    def render():
        v.Render(G.SHADE_TRILINEAR, 0, 0)

    ln0 = splat()
    # the autotuner on this workload: gather threads per brick, then the
    # march's ray-packet shape x samples fetched ahead
    tg = v.TuneGather(4, n, r, z, ln0)
    v.UpdateApron(0, 0.0)
    tr = v.TuneRenderShape(G.SHADE_TRILINEAR)
    tuned = dict(gather_threads=tg[0], gather_ms={str(k): x for k, x in tg[1].items()},
                 render_unroll=tr[0][0], render_tile="%dx%d" % (tr[0][1], 64 // tr[0][1]),
                 render_ms={"%dx%d/u%d" % (tw, 64 // tw, u): x for (tw, u), x in tr[1].items()})
    for _ in range(args.warmup):
        splat()
        render()
    v.synchronize()
    t_splat, t_render, kg, ka, kr, st = [], [], [], [], [], []
    entries = 0
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        entries = splat()
        v.synchronize()
        t1 = time.perf_counter()
        render()
        v.synchronize()
        t2 = time.perf_counter()
        t_splat.append(t1 - t0)
        t_render.append(t2 - t1)
        k = v.kernel_ms()
        kg.append(k["gather"]); ka.append(k["apron"]); kr.append(k["render"])
        st.append(v.stage_ms())
    v.synchronize()
    t_all = time.perf_counter() - t_all0
    # host-buffer callers: the points cross PCIe first (pageable host memory)
    t_up = []
    for _ in range(3):
        t0 = time.perf_counter()
        v.SetPoints(pts)
        v.synchronize()
        t_up.append(time.perf_counter() - t0)
    ab = algorithmic_bytes(v, n, entries)
    gather_ms = float(np.median(kg))
    apron_ms = float(np.median(ka))
    render_ms = float(np.median(kr))
    stages = {k: float(np.median([s[k] for s in st])) for k in st[0]}
    out = dict(n=n, entries=entries, t_splat=float(np.median(t_splat)), t_render=float(np.median(t_render)), t_all=t_all,
               t_upload=float(np.median(t_up)), autotune=tuned,
               ab=ab,
               gather_ms=gather_ms, apron_ms=apron_ms, render_ms=render_ms, stages=stages,
               rays=sc["width"] * sc["height"], topo=[len(v.pool(0, l)) for l in range(5)])
    v.close()
    return out


def _grid_scene(pts, w):
    """Emission-absorption view of one grid: camera on the diagonal of its
    bounding box at 2.2x the box size."""
    from gvdb_amd import workloads as W
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    c = tuple((lo + hi) / 2)
    return W.ea_scene(w, w, centre=c, dist=float(2.2 * np.max(hi - lo)))


def run_serving(args, grids):
    """Serving: the serving workload's grids (up to 200, initial points
    log-uniform in [50k, 4M]) stay resident; a stream of requests (70% renders of 256-1024 pixels a side from a random camera
    on a sphere of 1.5x the grid's extent, 30% inserts of 20k-1M points
    near the grid's existing points, each followed by the grid's rebuild)
    is served plainly (one request at a time, one-grid calls) and with
    continuous batching.  Saturation throughput comes from all requests
    arriving at once; latency from Poisson arrivals at 80% of the plain
    form's throughput."""
    import gvdb_amd as G
    from gvdb_amd import workloads as W
    g = np.random.default_rng(5)
    n_req = args.serve_requests
    S = G.Server(0, 250 << 30, 64)
    vols = []
    for kind, pts in grids:
        v = G.VolumeGVDB(0)
        v.Configure(5, 4, 3, 3, 3)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, G.T_FLOAT, 1)
        v.apply_scene(W.ea_scene(256, 256))
        S.add_grid(v, pts, 2.0)
        vols.append(v)
    plan = []
    for i in range(n_req):
        k = int(g.integers(0, len(grids)))
        if g.random() < 0.7:
            side = int(g.integers(32, 129)) * 8
            plan.append(("render", k, side, (float(g.uniform(-180, 180)), float(g.uniform(-80, 80)), 0.0)))
        else:
            m = int(round(np.exp(np.log(2e4) + (np.log(1e6) - np.log(2e4)) * g.random())))
            base = grids[k][1][g.integers(0, len(grids[k][1]), m)]
            pts = np.maximum(base + g.normal(0, 2.0, base.shape), 0.5).astype(np.float32)
            plan.append(("insert", k, pts, None))

    def submit(arrivals):
        S.clear_requests()
        for (kind, k, a, b), t in zip(plan, arrivals):
            if kind == "render":
                S.submit_render(k, t, a, a, G.SHADE_VOLUME, b, 0.0)
            else:
                S.submit_insert(k, t, a)

    def serve(mode, arrivals):
        S.reset()
        submit(arrivals)
        lat, st, wall = S.run(mode)
        assert np.all(st == 0), "requests failed: %s" % np.unique(st)
        return lat, wall

    zero = np.zeros(n_req)
    serve(S.CONTINUOUS, zero)          # warm-up
    lp0, wp0 = serve(S.PLAIN, zero)
    lc0, wc0 = serve(S.CONTINUOUS, zero)
    rate = 0.8 * n_req / wp0
    arrivals = np.cumsum(g.exponential(1.0 / rate, n_req))
    arrivals -= arrivals[0]
    lp, wp = serve(S.PLAIN, arrivals)
    lc, wc = serve(S.CONTINUOUS, arrivals)
    out = dict(workload="%d resident grids, initial points log-uniform in [50k, 4M] (%d total), %d requests: 70%% EA "
                        "renders 256-1024 px a side, 30%% inserts of 20k-1M points + rebuild"
                        % (len(grids), sum(len(p) for _, p in grids), n_req),
               requests=n_req, renders=sum(1 for p in plan if p[0] == "render"),
               saturation=dict(plain_requests_per_s=n_req / wp0, continuous_requests_per_s=n_req / wc0,
                               speedup=wp0 / wc0),
               poisson=dict(rate_per_s=rate,
                            plain=dict(requests_per_s=n_req / wp, p50_ms=1e3 * float(np.percentile(lp, 50)),
                                       p99_ms=1e3 * float(np.percentile(lp, 99))),
                            continuous=dict(requests_per_s=n_req / wc, p50_ms=1e3 * float(np.percentile(lc, 50)),
                                            p99_ms=1e3 * float(np.percentile(lc, 99)))))
    S.close()
    for v in vols:
        v.close()
    return out


def run_batched(args, grids):
    """The batched workload: 64 grids (point counts log-uniform in [10k, 2M],
    shell / slab / value-noise cloud), each splatted (topology, atlas, insert,
    gather, apron) and rendered at 512x512 (emission-absorption).  One request
    = one grid's splat + render.  Plain form: the grids one after another
    through the one-grid API.  Batched form: every stage once for all 64 grids
    (one launch per kernel).  Latency = completion time of a request, all 64
    submitted at once."""
    import gvdb_amd as G
    from gvdb_amd import workloads as W
    w = args.batch_width
    vols = []
    for kind, pts in grids:
        v = G.VolumeGVDB(0)
        v.Configure(5, 4, 3, 3, 3)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, G.T_FLOAT, 1)
        v.SetPoints(pts)
        v.apply_scene(_grid_scene(pts, w))
        v.AddRenderBuf(0, w, w, 4)
        vols.append(v)
    ns = [len(p) for _, p in grids]
    r, z = 2.0, (0.0, 0.0, 0.0)

    def plain():
        t0 = time.perf_counter()
        done = []
        for v, n in zip(vols, ns):
            v.RequestFullRebuild(True)
            v.RebuildTopology(n, r, z)
            v.FinishTopology(False, True)
            v.UpdateAtlas()
            v.ClearChannel(0)
            ln = v.InsertPointsSubcell(4, n, r, z)
            v.GatherDensity(4, n, r, z, ln, 0, 0, False)
            v.UpdateApron(0, 0.0)
            v.Render(G.SHADE_VOLUME, 0, 0)
            v.synchronize()
            done.append(time.perf_counter() - t0)
        return done

    B = G.GridBatch(0, len(vols))

    def batched():
        t0 = time.perf_counter()
        for v in vols:
            v.RequestFullRebuild(True)
        B.RebuildTopology(vols, ns, r)
        B.FinishTopology(vols, False, True)
        B.UpdateAtlas(vols)
        B.ClearChannel(vols, 0)
        lens = B.InsertPointsSubcell(vols, 4, ns, r)
        B.GatherDensity(vols, 4, lens, r, 0)
        B.UpdateApron(vols, 0, 0.0)
        B.Render(vols, G.SHADE_VOLUME, 0, 0)
        B.synchronize()
        t = time.perf_counter() - t0
        return [t] * len(vols)

    plain()
    batched()
    # the autotuner over the whole batch: gather threads per brick, apron
    # bricks per workgroup, ray-packet shape x samples fetched ahead
    lens = B.InsertPointsSubcell(vols, 4, ns, r)
    tg = B.Tune(vols, "gather", 4, lens, r)
    ta = B.Tune(vols, "apron")
    tr = B.Tune(vols, "render", shading=G.SHADE_VOLUME)
    tuned = dict(gather_threads=tg[0], gather_ms={str(k): x for k, x in tg[1].items()},
                 apron_run=ta[0], apron_ms={str(k): x for k, x in ta[1].items()},
                 render_unroll=tr[0][0], render_tile="%dx%d" % (tr[0][1], 64 // tr[0][1]),
                 render_ms={"%dx%d/u%d" % (tw, 64 // tw, u): x for (tw, u), x in tr[1].items()})
    for _ in range(max(1, args.warmup)):
        plain()
        batched()
    reps = max(1, min(args.steps, 5))
    lp, lb = [], []
    for _ in range(reps):
        lp.append(plain())
        lb.append(batched())
    # the same pixels either way
    img_b = [v.ReadRenderBuf(0, w, w, 4) for v in vols[:4]]
    plain()
    same = all(np.array_equal(v.ReadRenderBuf(0, w, w, 4), ib) for v, ib in zip(vols[:4], img_b))
    tp = float(np.median([max(x) for x in lp]))
    tb = float(np.median([max(x) for x in lb]))
    allp, allb = np.concatenate(lp), np.concatenate(lb)
    for v in vols:
        v.close()
    B.close()
    return dict(workload="64 grids, points log-uniform in [10k, 2M] (%d total), shell/slab/noise cloud, radius 2, "
                         "subcell 4, %dx%d emission-absorption render each" % (sum(ns), w, w),
                grids=len(ns), points=int(sum(ns)), reps=reps,
                plain=dict(ms=1e3 * tp, requests_per_s=len(ns) / tp, p50_ms=1e3 * float(np.percentile(allp, 50)),
                           p99_ms=1e3 * float(np.percentile(allp, 99))),
                batched=dict(ms=1e3 * tb, requests_per_s=len(ns) / tb, p50_ms=1e3 * float(np.percentile(allb, 50)),
                             p99_ms=1e3 * float(np.percentile(allb, 99))),
                speedup=tp / tb, same_images=bool(same), autotune=tuned)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(runs=5, render_w=1024):
    """The oracle's host path, timed on this host's cores: host ActivateSpace
    topology (single-threaded by construction), FinishTopology, UpdateAtlas,
    ClearChannel, subcell insert, FP32 gather and apron (OpenMP over
    subcells / bricks), and the FP32 isosurface march (OpenMP over pixels).
    All cores (OMP_NUM_THREADS, or the OpenMP default): the whole workload.
    One thread: the slab's first eighth in x (x < 106, with its plane strip),
    so every stage scales down together.  One warm-up, then `runs` runs;
    per-stage medians."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from gvdb_amd import workloads as W
    full = np.ascontiguousarray(W.slab_plane())
    crop = np.ascontiguousarray(full[full[:, 0] < 106.0])
    sc = W.iso_scene(render_w, render_w)
    blk = O.scene_block(sc)
    tf = O.linear_tf(*sc["tf"])
    names = ("topology", "atlas", "insert", "gather", "apron")

    def leg(pts, threads):
        O.set_threads(threads)
        splat, stages, render = [], [], []
        for k in range(runs + 1):
            o = O.Oracle(fp64=False)
            t0 = time.perf_counter()
            o.rebuild_topology_cpu(pts)
            o.finish_topology()
            o.update_atlas()
            o.clear_channel()
            o.insert_points_subcell(pts, 2.0)
            o.gather_density(2.0, 4, 32)
            o.update_apron(0.0, 32)
            t1 = time.perf_counter()
            o.render(blk, tf, 4, 32)
            t2 = time.perf_counter()
            if k:
                splat.append(t1 - t0)
                render.append(t2 - t1)
                stages.append([o.stage_ms(i) for i in range(5)])
            del o
        st = np.median(np.asarray(stages), axis=0)
        ts, tr = float(np.median(splat)), float(np.median(render))
        return dict(points=len(pts), threads=O.max_threads(), splat_s=ts, render_s=tr, points_per_s=len(pts) / ts,
                    mrays_per_s=render_w * render_w / tr / 1e6, stage_ms={n: float(x) for n, x in zip(names, st)})

    allc = leg(full, int(os.environ.get("OMP_NUM_THREADS", "0")) or O.max_threads())
    one = leg(crop, 1)
    return dict(all_cores=allc, one_thread=one, runs=runs, cpu_model=cpu_model(), nproc=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=("all cores: the whole workload (%d points) and the %dx%d isosurface, median of %d runs after "
                        "one warm-up; one thread: the points with x < 106 (%d points) and the same render"
                        % (len(full), render_w, render_w, runs, len(crop))))


def stream_copy_gbs(mib=2048, reps=10):
    """Achievable HBM bandwidth: a device-to-device copy (the runtime's copy
    kernel) of `mib` MiB timed with HIP events; bytes moved = 2 x size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    best = []
    for k in range(reps + 2):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None) == 0     # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float(0)
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        if k >= 2:
            best.append(ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * n / (float(np.median(best)) * 1e-3) / 1e9


def progress(msg):
    """One line per finished part on stderr (the JSON result goes to stdout)."""
    print("[bench %.0fs] %s" % (time.time() - T_START, msg), file=sys.stderr, flush=True)


def copy_calibration(mbytes):
    """One device-to-device hipMemcpy of `mbytes` MiB (a blit kernel moving a
    known number of bytes), for scaling the FETCH_SIZE / WRITE_SIZE counters
    of a profiled run."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mbytes) << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps(dict(copy_calibration_bytes=n, copies=3)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-grid batched comparison")
    ap.add_argument("--batch-width", type=int, default=512)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving comparison")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-grids", type=int, default=200, help="resident grids of the serving workload")
    ap.add_argument("--no-stream", action="store_true", help="skip the copy-bandwidth measurement")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the point-to-volume path runs one grid per GPU and this build has "
              "no multi-GPU form (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    g = run_gpu(args)
    progress("medium workload timed")
    ab = g["ab"]
    gather_gbs = ab["gather"] / (g["gather_ms"] * 1e-3) / 1e9
    apron_gbs = ab["apron"] / (g["apron_ms"] * 1e-3) / 1e9
    value = g["n"] / g["t_splat"]
    line = dict(
        metric=METRIC, value=value, unit="points/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * g["t_all"] / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32", data="synthetic",
        config=dict(workload="plain medium: 4,194,304 slab + 200,000 plane points, tree {3,3,3,4,5}, radius 2, "
                             "subcell 4, isosurface 0.5 trilinear %dx%d step 0.25" % (args.width, args.width),
                    points=g["n"], subcell_entries=g["entries"], bricks=ab["B"], nodes_per_level=g["topo"]),
        mrays_per_s=g["rays"] / g["t_render"] / 1e6,
        splat_ms=1e3 * g["t_splat"], render_ms=1e3 * g["t_render"], stage_ms=g["stages"],
        upload_ms=1e3 * g["t_upload"], pcie_inclusive_points_per_s=g["n"] / (g["t_splat"] + g["t_upload"]),
        autotune=g["autotune"],
        kernel_ms=dict(gather=g["gather_ms"], apron=g["apron_ms"], render=g["render_ms"]),
        roofline=dict(bound="hbm", kernel="gather", achieved=gather_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=gather_gbs / HBM_PEAK_GBS, traffic=profiled_traffic("k_push_gather<128, float, 4, false>"),
                      algorithmic_bytes=ab["gather"], bytes_formula="E*12 + S*8 + B*512*4",
                      design_bytes=ab["gather_design"], design_frac=ab["gather_design"] / (g["gather_ms"] * 1e-3) / 1e9 /
                      HBM_PEAK_GBS),
        roofline_apron=dict(bound="hbm", kernel="apron", achieved=apron_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                            frac=apron_gbs / HBM_PEAK_GBS, traffic=profiled_traffic("k_apron_flat<float, 8>"),
                            algorithmic_bytes=ab["apron"], bytes_formula="B*488*8"),
        batched=None,
    )
    if not args.no_stream:
        cp = stream_copy_gbs()
        progress("copy bandwidth measured")
        line["copy_bandwidth"] = dict(achieved=cp, peak=HBM_PEAK_GBS, unit="GB/s", frac=cp / HBM_PEAK_GBS,
                                      kernel="device-to-device hipMemcpyAsync, 2 GiB, median of 10",
                                      gather_frac_of_copy=gather_gbs / cp, apron_frac_of_copy=apron_gbs / cp)
    if not args.no_batched:
        from gvdb_amd import workloads as W
        line["batched"] = run_batched(args, W.batch_grids())
        progress("batched workload timed")
    if not args.no_serve:
        from gvdb_amd import workloads as W
        line["serving"] = run_serving(args, W.serve_grids(count=args.serve_grids))
        progress("serving workload timed")
    if not args.no_cpu:
        c = run_cpu_baseline()
        progress("CPU baseline timed")
        a = c["all_cores"]
        line["cpu_baseline"] = dict(value=a["points_per_s"], unit="points/s", cores=a["threads"], kind="port",
                                    sample=c["sample"], mrays_per_s=a["mrays_per_s"], splat_s=a["splat_s"],
                                    render_s=a["render_s"], stage_ms=a["stage_ms"], one_thread=c["one_thread"],
                                    cpu_model=c["cpu_model"], nproc=c["nproc"], omp_num_threads=c["omp_num_threads"],
                                    runs=c["runs"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
