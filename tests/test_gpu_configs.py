"""Parity at the benchmark configurations' own sizes.

Each test runs one benchmark configuration at full size through the C ABI and
compares it with the CPU oracle on the same seeded inputs:

  small      262,144 shell points, host topology build, 256x256 emission-
             absorption march at step 0.5 (density scale 0.05);
  medium     4,194,304 slab + 200,000 plane points, GPU topology, 1024x1024
             isosurface at 0.5, step 0.25;
  batched    the 64-grid workload (10k-2M points, shell/slab/cloud) through the
             batched calls, 512x512 emission-absorption renders;
  precision  16.4M points in 32 Gaussian clusters, FP32 and FP16 atlases,
             2048x2048 emission-absorption render;
  serving    200 resident grids with 50k-4M points each, a request mix served
             one request at a time and with continuous batching.

Bars: topology, atlas slots, subcell entry counts bit-exact; densities within
1e-5 relative of the FP64 gather with no nonzero where FP64 is exactly zero
(FP16 atlas: within the half rounding bound 2^-11 |b| + 2^-25); images within
1e-3 RMSE of the oracle march on float RGBA; the uchar4 buffer equal to the
float buffer converted as the kernel stores it (x * 255, truncated).
"""
import time

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
RMSE_TOL = 1e-3
_T0 = time.perf_counter()


def log(msg):
    """Progress of the long checks (shown with pytest -s)."""
    print("[%7.1fs] %s" % (time.perf_counter() - _T0, msg), flush=True)


def _render_both(gpu, v, sc, shading):
    """The float RGBA image and the uchar4 image of one scene."""
    w, h = sc["width"], sc["height"]
    v.apply_scene(sc)
    v.AddRenderBuf(0, w, h, 16)
    v.AddRenderBuf(1, w, h, 4)
    v.Render(shading, 0, 0)
    v.Render(shading, 0, 1)
    return v.ReadRenderBuf(0, w, h, 16), v.ReadRenderBuf(1, w, h, 4)


def _uchar_of(img):
    return (img.astype(np.float32) * np.float32(255.0)).astype(np.uint8)


def _check_atlas(v, o, ref):
    info = o.atlas_info()
    assert v.atlas_info() == info
    amap = o.atlas_map()
    assert np.array_equal(v.atlas_map().view(np.int32), amap.view(np.int32))
    dmax, rel, zeros = H.brick_density_errors(v.atlas(), ref, amap, info["cnt"])
    assert zeros == 0, "GPU density nonzero where the FP64 oracle is exactly zero"
    assert rel <= REL_TOL, "max relative density error %.3g (abs %.3g)" % (rel, dmax)


def test_small_config_host_topology(gpu):
    """Plain small: 262,144 points on a sphere shell (radius 120 + N(0, 1.5)
    around (256,256,256)), host ActivateSpace topology, radius 2, 256x256
    emission-absorption at step 0.5."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.sphere_shell()
    assert len(pts) == 262144
    log("small: %d points" % len(pts))
    v, n = H.gpu_volume(gpu, pts, 2.0, host_topology=True)
    o, m = H.oracle_volume(O, pts, 2.0, host_topology=True, precision=64)
    assert n == m
    H.assert_pools_equal(v, o)
    _check_atlas(v, o, o.atlas(64))
    # the GPU build at r < 1 makes the same node set (in its own pool order)
    g, _ = H.gpu_volume(gpu, pts, 0.0, stop_after="topology")
    assert H.canonical_nodes(g) == H.canonical_nodes(v)
    g.close()
    sc = W.ea_scene(256, 256)
    assert sc["steps"][0] == 0.5
    img, img8 = _render_both(gpu, v, sc, gpu.SHADE_VOLUME)
    ref, capped = o.render(v.scene_block(256, 256), O.linear_tf(*sc["tf"]), 7, 64, capped=True)
    assert not capped.any()
    assert H.rmse(img, ref) <= RMSE_TOL
    assert np.array_equal(img8, _uchar_of(img))
    v.close()


def test_medium_config(gpu):
    """Plain medium: 4,194,304 slab points + 200,000 plane points, GPU
    topology, radius 2, 1024x1024 isosurface at 0.5 with step 0.25."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.slab_plane()
    assert len(pts) == 4394304
    log("medium: %d points" % len(pts))
    v, n = H.gpu_volume(gpu, pts, 2.0)
    assert n == 49790999
    log("medium: GPU splat done")
    o, m = H.oracle_volume(O, pts, 2.0, precision="both")
    assert m == n and o.undefined_events() == 0
    H.assert_pools_equal(v, o)
    vi, ob = v.info(), o.bounds()
    assert vi["top_lev"] == o.top_lev()
    for k in ("obj_min", "obj_max", "pos_min", "pos_max"):
        assert np.array_equal(vi[k], ob[k]), k
    _check_atlas(v, o, o.atlas(64))
    log("medium: pools and densities checked")
    sc = W.iso_scene(1024, 1024)
    assert sc["steps"][0] == 0.25 and sc["vrange"][0] == 0.5
    img, img8 = _render_both(gpu, v, sc, gpu.SHADE_TRILINEAR)
    blk, tf = v.scene_block(1024, 1024), O.linear_tf(*sc["tf"])

    def flips(ref):
        return int(np.count_nonzero(np.abs(img.astype(np.float64) - ref).max(-1) > 1e-3))

    # The isosurface is a threshold test (first sample >= 0.5, then a shadow
    # ray), so a sample within a gather's rounding of 0.5 can end a ray one
    # step apart or flip its shadow: end to end such rays are counted.
    # (1) the oracle's FP32 gather (the reference's sqrtf formula, ~1e-7
    # absolute) and FP32 march: reported
    ref, capped = o.render(blk, tf, 4, 32, capped=True)
    assert not capped.any()
    log("medium: against the oracle's FP32 gather + march: RMSE %.3g, %d rays differ by > 1e-3"
        % (H.rmse(img, ref), flips(ref)))
    # (2) the FP64 gather rounded to FP32, marched in FP32: at most 1 ray in 10^4
    o.set_atlas(o.atlas(64).astype(np.float32), 32)
    ref, capped = o.render(blk, tf, 4, 32, capped=True)
    assert not capped.any()
    f = flips(ref)
    log("medium: against the FP64 gather + FP32 march: RMSE %.3g, %d rays differ by > 1e-3" % (H.rmse(img, ref), f))
    assert f <= 1024 * 1024 // 10000, "%d rays differ end to end" % f
    # (3) the march itself: the oracle's FP32 march over the GPU's atlas
    o.set_atlas(v.atlas(), 32)
    ref, capped = o.render(blk, tf, 4, 32, capped=True)
    assert not capped.any()
    e = H.rmse(img, ref)
    log("medium: march on the same atlas: RMSE %.3g, %d rays differ by > 1e-3" % (e, flips(ref)))
    assert e <= RMSE_TOL, "image RMSE %.3g against the oracle's FP32 march" % e
    assert np.array_equal(img8, _uchar_of(img))
    v.close()


def _grid_scene(pts, w):
    from gvdb_amd import workloads as W
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    return W.ea_scene(w, w, centre=tuple((lo + hi) / 2), dist=float(2.2 * np.max(hi - lo)))


def test_batched_config_64_grids(gpu):
    """Batched: the 64-grid workload (point counts log-uniform in [10k, 2M];
    shell, slab, value-noise cloud), every stage one batched call for all
    grids, 512x512 emission-absorption renders.  Every grid against the
    oracle (pools, entry count, densities, image) and its batched image
    equal to the one-grid calls' image."""
    import oracle as O
    from gvdb_amd import workloads as W
    grids = W.batch_grids()
    assert len(grids) == 64
    w = 512
    vols = []
    for _, pts in grids:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.SetPoints(pts)
        v.apply_scene(_grid_scene(pts, w))
        v.AddRenderBuf(0, w, w, 16)
        vols.append(v)
    ns = [len(p) for _, p in grids]
    B = gpu.GridBatch(0, 64)
    for v in vols:
        v.RequestFullRebuild(True)
    B.RebuildTopology(vols, ns, 2.0)
    B.FinishTopology(vols, False, True)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols, 0)
    lens = B.InsertPointsSubcell(vols, 4, ns, 2.0)
    B.GatherDensity(vols, 4, lens, 2.0, 0)
    B.UpdateApron(vols, 0, 0.0)
    B.Render(vols, gpu.SHADE_VOLUME, 0, 0)
    B.synchronize()
    batched = [v.ReadRenderBuf(0, w, w, 16) for v in vols]
    log("batched: 64 grids splatted and rendered")
    capped_total = 0
    for k, ((kind, pts), v) in enumerate(zip(grids, vols)):
        o, m = H.oracle_volume(O, pts, 2.0, precision=64)
        assert lens[k] == m, (k, kind)
        H.assert_pools_equal(v, o)
        _check_atlas(v, o, o.atlas(64))
        ref, capped = o.render(v.scene_block(w, w), O.linear_tf(*_grid_scene(pts, w)["tf"]), 7, 64, capped=True)
        capped_total += int(capped.sum())
        keep = ~capped
        e = H.rmse(batched[k][keep], ref[keep])
        assert e <= RMSE_TOL, "grid %d (%s): image RMSE %.3g" % (k, kind, e)
        log("batched: grid %d (%s, %d points) matches the oracle" % (k, kind, len(pts)))
        del o, ref
    assert capped_total == 0
    # the one-grid calls give the same pixels
    for k, (v, n) in enumerate(zip(vols, ns)):
        v.RequestFullRebuild(True)
        v.RebuildTopology(n, 2.0, H.ZERO3)
        v.FinishTopology(False, True)
        v.UpdateAtlas()
        v.ClearChannel(0)
        ln = v.InsertPointsSubcell(4, n, 2.0, H.ZERO3)
        v.GatherDensity(4, n, 2.0, H.ZERO3, ln, 0, 0, False)
        v.UpdateApron(0, 0.0)
        v.Render(gpu.SHADE_VOLUME, 0, 0)
        assert np.array_equal(v.ReadRenderBuf(0, w, w, 16), batched[k]), k
    B.close()
    for v in vols:
        v.close()


def test_precision_config_fp16_fp32(gpu):
    """Precision: 32 Gaussian clusters (sigma 20-80, centres in 2048^3) of
    16,777,216 drawn points (those inside [0, 2048)^3 kept), FP32 and FP16
    atlases with FP32 arithmetic, 2048x2048 emission-absorption render;
    each against the FP64 oracle."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.clusters()
    assert 16_000_000 < len(pts) <= 16777216
    width = 2048
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    sc = W.ea_scene(width, width, centre=tuple((lo + hi) / 2), dist=float(1.6 * np.max(hi - lo)))
    res = {}
    for name, t in (("fp32", gpu.T_FLOAT), ("fp16", gpu.T_HALF)):
        v, n = H.gpu_volume(gpu, pts, 2.0, atlas_type=t)
        v.apply_scene(sc)
        v.AddRenderBuf(0, width, width, 16)
        v.Render(gpu.SHADE_VOLUME, 0, 0)
        res[name] = dict(img=v.ReadRenderBuf(0, width, width, 16), n=n, blk=v.scene_block(width, width),
                         atlas=v.atlas(0, np.float32 if t == gpu.T_FLOAT else np.float16),
                         pools=[v.pool(0, l) for l in range(5)])
        v.close()
        log("precision: GPU %s done" % name)
    assert res["fp32"]["n"] == res["fp16"]["n"]
    o, m = H.oracle_volume(O, pts, 2.0, precision=64)
    log("precision: oracle FP64 gather done")
    assert m == res["fp32"]["n"]
    for lev in range(5):
        for f in H.NODE_FIELDS:
            assert np.array_equal(res["fp32"]["pools"][lev][f], o.pool(0, lev)[f]), (lev, f)
    info = o.atlas_info()
    amap = o.atlas_map()
    ref = o.atlas(64)
    dmax, rel, zeros = H.brick_density_errors(res["fp32"]["atlas"], ref, amap, info["cnt"])
    assert zeros == 0 and rel <= REL_TOL, "FP32 atlas: max rel %.3g" % rel
    worst, zeros16 = H.half_errors(res["fp16"]["atlas"], ref, amap, info["cnt"])
    assert zeros16 == 0 and worst <= 1.0, "FP16 atlas: error %.3g of the half rounding bound" % worst
    del ref, res["fp32"]["atlas"], res["fp16"]["atlas"]
    log("precision: atlases checked")
    tf = O.linear_tf(*sc["tf"])
    r64, cap64 = o.render(res["fp32"]["blk"], tf, 7, 64, capped=True)
    o.gather_density(2.0, 4, 32)
    o.update_apron(0.0, 32)
    r32, cap32 = o.render(res["fp32"]["blk"], tf, 7, 32, capped=True)
    log("precision: oracle marches done")
    # the parity claim: the reference's FP32 march, every ray
    for name in ("fp32", "fp16"):
        e = H.rmse(res[name]["img"], r32)
        assert e <= RMSE_TOL, "%s atlas: RMSE %.3g against the oracle FP32 march" % (name, e)
    # against FP64 on the rays where the FP32 and FP64 walks agree (within the
    # walk cap in both and ending within 1e-3 of each other)
    agree = ~(cap64 | cap32) & (np.abs(r32.astype(np.float64) - r64).max(-1) <= 1e-3)
    assert agree.mean() > 0.5
    for name in ("fp32", "fp16"):
        e = H.rmse(res[name]["img"][agree], r64[agree])
        assert e <= RMSE_TOL, "%s atlas: RMSE %.3g against the oracle FP64 march" % (name, e)


def test_serving_config_200_grids(gpu):
    """Serving: 200 resident grids with initial point counts log-uniform in
    [50k, 4M]; a Poisson-timed mix of 70% renders (256-1024 pixels a side,
    camera on a sphere of 1.5x the grid's extent) and 30% inserts (20k-1M
    points near the grid's points, each followed by the grid's rebuild).
    Every render gives the same pixels served alone (one request at a time)
    and in the mix (continuous batching), and a grid that took inserts ends
    as if built from all its points."""
    from gvdb_amd import workloads as W
    grids = W.serve_grids()
    assert len(grids) == 200
    S = gpu.Server(0, 250 << 30, 64)
    vols = []
    for _, pts in grids:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.apply_scene(W.ea_scene(256, 256))
        S.add_grid(v, pts, 2.0)
        vols.append(v)
    log("serving: %d grids, %d points" % (len(grids), sum(len(p) for _, p in grids)))
    S.reset()
    log("serving: grids built")
    g = np.random.default_rng(55)
    n_req = 60
    arrivals = np.cumsum(g.exponential(1.0 / 400.0, n_req))
    inserts = {}
    for t in arrivals:
        k = int(g.integers(0, len(grids)))
        if g.random() < 0.7:
            side = int(g.integers(32, 129)) * 8
            S.submit_render(k, t - arrivals[0], side, side,
                            angs=(float(g.uniform(-180, 180)), float(g.uniform(-80, 80)), 0.0))
        else:
            m = int(round(np.exp(np.log(2e4) + (np.log(1e6) - np.log(2e4)) * g.random())))
            base = grids[k][1][g.integers(0, len(grids[k][1]), m)]
            pts = np.maximum(base + g.normal(0, 2.0, base.shape), 0.5).astype(np.float32)
            S.submit_insert(k, t - arrivals[0], pts)
            inserts.setdefault(k, []).append(pts)
    lat0, st0, _ = S.run(S.PLAIN)
    log("serving: plain pass done")
    assert np.all(st0 == 0)
    alone = [r[2].copy() for r in S.requests if r[0] == "render"]
    S.reset()
    lat1, st1, _ = S.run(S.CONTINUOUS)
    assert np.all(st1 == 0)
    mixed = [r[2] for r in S.requests if r[0] == "render"]
    assert len(alone) == len(mixed) > 20
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)
    for k in sorted(inserts)[:4]:
        pts = np.concatenate([grids[k][1]] + inserts[k])
        ref, _ = H.gpu_volume(gpu, pts, 2.0)
        H.assert_pools_equal(vols[k], ref)
        assert np.array_equal(vols[k].atlas(), ref.atlas())
        ref.close()
    S.close()
    for v in vols:
        v.close()
