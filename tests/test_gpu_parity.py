"""Parity of the HIP path (through the C ABI) with the CPU oracle on identical
seeded inputs.

Bars (stated per test):
  topology, atlas slots, subcell counts/offsets/positions: bit-exact;
  subcell entries: bit-exact after sorting inside each subcell (the
  reference appends with atomics, so order inside a subcell is unspecified);
  densities: within 1e-5 relative of the FP64 oracle gather;
  images: RMSE <= 1e-3 against the FP64 oracle march on a float RGBA buffer.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _sphere(n, seed=11):
    from gvdb_amd import workloads as W
    return W.sphere_shell(n, seed=seed)


def _slab(n, seed=12):
    from gvdb_amd import workloads as W
    return W.slab_plane(n, n // 20, seed=seed, lo=(40.0, 30.0, 40.0), hi=(200.0, 90.0, 160.0), plane_y=20.0)


CASES = {
    "sphere_r2": (lambda: _sphere(20000), 2.0),
    "sphere_r0.5": (lambda: _sphere(5000, 13), 0.5),
    "sphere_r1.5": (lambda: _sphere(5000, 14), 1.5),
    "slab_r2": (lambda: _slab(30000), 2.0),
    "tiny_r2": (lambda: (np.random.default_rng(5).random((200, 3)) * 40 + 70).astype(np.float32), 2.0),
    # bricks at the origin, where coordinates are finer than any fixed-point scale
    "origin_r2": (lambda: np.concatenate([np.random.default_rng(6).random((3000, 3)) * 24,
                                          [[0.0, 0.0, 0.0], [0.0, 5.5, 1e-3], [3e-5, 0.25, 9.0]]]).astype(np.float32),
                  2.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_topology_bit_exact(gpu, case):
    import oracle as O
    pts, r = CASES[case][0](), CASES[case][1]
    v, _ = H.gpu_volume(gpu, pts, r, stop_after="topology")
    o, _ = H.oracle_volume(O, pts, r, stop_after="topology")
    assert o.undefined_events() == 0
    H.assert_pools_equal(v, o)
    vi = v.info()
    assert vi["top_lev"] == o.top_lev()
    ob = o.bounds()
    assert np.array_equal(vi["obj_min"], ob["obj_min"]) and np.array_equal(vi["obj_max"], ob["obj_max"])
    assert np.array_equal(vi["pos_min"], ob["pos_min"]) and np.array_equal(vi["pos_max"], ob["pos_max"])


@pytest.mark.parametrize("case", ["sphere_r2", "slab_r2", "tiny_r2", "sphere_r0.5", "origin_r2"])
def test_atlas_and_subcells_bit_exact(gpu, case):
    import oracle as O
    pts, r = CASES[case][0](), CASES[case][1]
    v, n = H.gpu_volume(gpu, pts, r, stop_after="insert")
    o, m = H.oracle_volume(O, pts, r, stop_after="insert")
    assert n == m
    H.assert_pools_equal(v, o)          # mValue written by UpdateAtlas
    assert v.atlas_info() == o.atlas_info()
    assert np.array_equal(v.atlas_map().view(np.int32), o.atlas_map().view(np.int32))
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    assert np.array_equal(gs["pos"], os_["pos"])
    assert np.array_equal(gs["nid"], os_["nid"])
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))


@pytest.mark.parametrize("case", ["sphere_r2", "slab_r2", "tiny_r2", "sphere_r1.5", "origin_r2"])
def test_density_and_apron_vs_fp64(gpu, case):
    import oracle as O
    pts, r = CASES[case][0](), CASES[case][1]
    v, n = H.gpu_volume(gpu, pts, r)
    o, m = H.oracle_volume(O, pts, r, precision=64)
    assert n == m
    a = v.atlas()
    b = o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    dmax, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0, "GPU density nonzero where the FP64 oracle is exactly zero"
    assert rel <= REL_TOL, "max relative density error %.3g (abs %.3g)" % (rel, dmax)


@pytest.mark.parametrize("case", ["sphere_r2", "slab_r2"])
def test_apron_boundary_value(gpu, case):
    """UpdateApron(chan, boundval) with a nonzero boundary value: apron voxels
    with no neighbour brick take it, the others their neighbour's value."""
    import oracle as O
    pts, r = CASES[case][0](), CASES[case][1]
    v, n = H.gpu_volume(gpu, pts, r, boundval=0.375)
    o, m = H.oracle_volume(O, pts, r, precision=64, boundval=0.375)
    assert n == m
    a, b = v.atlas(), o.atlas(64)
    assert np.count_nonzero(b == 0.375) > 0
    _, rel, zeros = H.brick_density_errors(a, b, o.atlas_map(), o.atlas_info()["cnt"])
    assert zeros == 0 and rel <= REL_TOL, rel
    v.close()


@pytest.mark.parametrize("shading", [7, 4])
def test_render_vs_fp64(gpu, shading):
    import oracle as O
    from gvdb_amd import workloads as W
    pts = _sphere(20000)
    v, _ = H.gpu_volume(gpu, pts, 2.0)
    o, _ = H.oracle_volume(O, pts, 2.0, precision=64)
    sc = W.ea_scene(96, 80) if shading == 7 else W.iso_scene(96, 80, centre=(256.0, 256.0, 256.0), dist=600.0)
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 16)
    v.Render(shading, 0, 0)
    img = v.ReadRenderBuf(0, sc["width"], sc["height"], 16)
    blk = v.scene_block(sc["width"], sc["height"])
    ref, capped = o.render(blk, O.linear_tf(*sc["tf"]), shading, 64, capped=True)
    assert not capped.any()          # no ray reaches the MAX_ITER walk cap in this scene
    e = H.rmse(img, ref)
    assert e <= 1e-3, "image RMSE %.3g" % e
    assert np.abs(img - ref).max() < 0.5


@pytest.mark.parametrize("shading", [7, 4])
def test_render_unroll_same_image(gpu, shading):
    """Samples fetched ahead per brick step (1, 2, 4) change only the order of
    the loads: the images are bit-identical, and the autotuner keeps one of
    the three."""
    from gvdb_amd import workloads as W
    pts = _sphere(20000)
    v, _ = H.gpu_volume(gpu, pts, 2.0)
    sc = W.ea_scene(96, 80) if shading == 7 else W.iso_scene(96, 80, centre=(256.0, 256.0, 256.0), dist=600.0)
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 16)
    imgs = []
    try:
        for u in (1, 2, 4):
            gpu.VolumeGVDB.SetRenderUnroll(shading, u)
            v.Render(shading, 0, 0)
            imgs.append(v.ReadRenderBuf(0, sc["width"], sc["height"], 16))
        best, ms = v.TuneRender(shading)
        assert best in (1, 2, 4) and all(t > 0 for t in ms.values())
    finally:
        gpu.VolumeGVDB.SetRenderUnroll(shading, 4 if shading == 7 else 2)
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])
    assert imgs[0][..., 3].max() > 0


def test_render_fp32_walk_at_large_coordinates(gpu):
    """Clusters in a 2048^3 space seen from ~3000 voxels away: there the
    reference's 1e-3 epsilon is a few float ulps, and the FP32 walk leaves
    the tree on some rays where the FP64 walk goes on (291 of 25600 rays
    differ by more than 1e-3 between the oracle's two marches on this
    scene).  The GPU march repeats the FP32 operations without contraction,
    so it follows the oracle's FP32 walk ray for ray."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.clusters(300000, k=12, seed=8, space=2048.0, sig=(20.0, 60.0))
    v, _ = H.gpu_volume(gpu, pts, 2.0)
    o, _ = H.oracle_volume(O, pts, 2.0, precision=32)
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    sc = W.ea_scene(160, 160, centre=tuple((lo + hi) / 2), dist=float(1.6 * np.max(hi - lo)))
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 16)
    v.Render(7, 0, 0)
    img = v.ReadRenderBuf(0, sc["width"], sc["height"], 16)
    ref = o.render(v.scene_block(sc["width"], sc["height"]), O.linear_tf(*sc["tf"]), 7, 32)
    e = H.rmse(img, ref)
    assert e <= 1e-3, "image RMSE %.3g against the FP32 march" % e
    assert np.abs(img - ref).max() < 5e-3


def test_apron_follows_every_atlas_write(gpu):
    """The apron copies neighbour faces from the gather's compact face buffer
    only while nothing else has written the atlas.  After a ClearChannel, a
    gather into another channel, an accumulate gather or a second apron, each
    apron must equal what the brick interiors say (the oracle's apron of the
    same interiors)."""
    import oracle as O
    pts = _sphere(20000)
    n = len(pts)
    v, m = H.gpu_volume(gpu, pts, 2.0)
    v.AddChannel(1, gpu.T_FLOAT, 1)
    o, _ = H.oracle_volume(O, pts, 2.0, precision=64)
    mask = H.used_brick_mask(o.atlas(64).shape, o.atlas_map(), o.atlas_info()["cnt"])
    z = (0.0, 0.0, 0.0)

    def check(expect_zero):
        a, b = v.atlas(0), o.atlas(64)
        if expect_zero:
            assert not np.any(a[mask]), "apron kept densities after ClearChannel"
        else:
            _, rel, zeros = H.density_errors(a[mask], b[mask])
            assert zeros == 0 and rel <= REL_TOL
    check(False)                                   # gather -> apron (face buffer)
    v.UpdateApron(0, 0.0)                          # a second apron, same result
    check(False)
    v.ClearChannel(0)
    v.UpdateApron(0, 0.0)                          # faces are stale now: all zero
    check(True)
    v.GatherDensity(4, n, 2.0, z, m, 0, 0, False)
    v.GatherDensity(4, n, 2.0, z, m, 1, 0, False)  # channel 1 takes the face buffer
    v.UpdateApron(0, 0.0)
    check(False)
    v.ClearChannel(0)
    v.GatherDensity(4, n, 2.0, z, m, 0, 0, True)   # accumulate over zeros: same interiors
    v.UpdateApron(0, 0.0)
    check(False)
    v.close()


def test_host_topology_path(gpu):
    """RebuildTopologyCPU (host ActivateSpace) then the GPU stages."""
    import oracle as O
    pts = _sphere(8000, 21)
    v, n = H.gpu_volume(gpu, pts, 2.0, host_topology=True)
    o, m = H.oracle_volume(O, pts, 2.0, host_topology=True, precision=64)
    assert n == m
    H.assert_pools_equal(v, o)
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL


@pytest.mark.parametrize("variant", ["large_radius", "orphans", "subcell2", "radius3", "radius2.25"])
def test_insert_paths(gpu, variant):
    """Both binning paths against the oracle: the pull path (own-subcell
    buckets) and the push path it falls back to for points outside every brick
    (topology built from a subset), neighbourhoods wider than one brick, and
    subcell sizes with more than 32 subcells per brick; radii 2, 2.25 and 3
    take the scatter gather's three variants (four or five voxels per axis
    of a point's box, or any number)."""
    import oracle as O
    pts = _sphere(6000, 31)
    n_topo, r_topo, r_ins, scs = len(pts), 2.0, 2.0, 4
    if variant == "large_radius":
        r_topo = r_ins = 8.0
    elif variant == "orphans":
        n_topo = 3000
    elif variant == "subcell2":
        scs = 2
    elif variant == "radius3":
        r_topo = r_ins = 3.0
    elif variant == "radius2.25":             # five voxels per axis of a point's box
        r_topo = r_ins = 2.25
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RebuildTopology(n_topo, r_topo, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(scs, len(pts), r_ins, H.ZERO3)
    v.GatherDensity(scs, len(pts), r_ins, H.ZERO3, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    o = O.Oracle()
    assert o.rebuild_topology(pts[:n_topo], r_topo) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, m = o.insert_points_subcell(pts, r_ins, scs)
    assert rc == 0 and n == m
    o.gather_density(r_ins, scs, 64)
    o.update_apron(0.0, 64)
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL


@pytest.mark.parametrize("trans", [(-60.0, 5.0, 0.0), (20.0, 0.0, 0.0)])
def test_translated_points(gpu, trans):
    """RebuildTopology(orig) and InsertPointsSubcell/GatherDensity(trans) with
    a translation that moves some points below zero (dropped), or moves some
    negative points above zero (dropped in both binning passes: the reference
    tests them inconsistently there)."""
    import oracle as O
    pts = _sphere(6000, 51)
    extra = np.random.default_rng(52).random((300, 3)).astype(np.float32) * 30
    extra[:, 0] -= 15.0          # x in [-15, 15)
    pts = np.concatenate([pts, extra + np.float32(240.0) * np.array([0, 1, 1], np.float32)])
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RebuildTopology(len(pts), 2.0, trans)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(4, len(pts), 2.0, trans)
    v.GatherDensity(4, len(pts), 2.0, trans, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0, trans) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, m = o.insert_points_subcell(pts, 2.0, 4, trans)
    assert rc == 0 and n == m
    o.gather_density(2.0, 4, 64)
    o.update_apron(0.0, 64)
    H.assert_pools_equal(v, o)
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL


def test_fp16_channel_vs_fp64(gpu):
    """Optional FP16 atlas channel: the gather and apron compute in FP32 and
    store FP16; each voxel is within FP16 rounding of the FP64 oracle
    (|a - b| <= 2^-10 |b| + 2^-24), and the emission-absorption image from the
    FP16 atlas (sampled in FP32) stays within 2e-3 RMSE of the FP64 march."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.clusters(30000, k=4, seed=91, space=400.0, sig=(10.0, 30.0))
    v, n = H.gpu_volume(gpu, pts, 2.0, atlas_type=gpu.T_HALF)
    o, m = H.oracle_volume(O, pts, 2.0, precision=64)
    assert n == m
    H.assert_pools_equal(v, o)
    a = v.atlas(0, np.float16).astype(np.float64)
    b = o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    d = np.abs(a[mask] - b[mask])
    assert np.all(d <= 2.0 ** -10 * np.abs(b[mask]) + 2.0 ** -24)
    assert np.count_nonzero((b[mask] == 0) & (a[mask] != 0)) == 0
    sc = W.ea_scene(96, 80, centre=tuple(float(x) for x in pts.mean(0)), dist=500.0)
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 16)
    v.Render(gpu.SHADE_VOLUME, 0, 0)
    img = v.ReadRenderBuf(0, sc["width"], sc["height"], 16)
    ref = o.render(v.scene_block(sc["width"], sc["height"]), O.linear_tf(*sc["tf"]), 7, 64)
    assert H.rmse(img, ref) <= 2e-3
