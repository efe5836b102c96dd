"""The batched form: one launch per kernel over many grids of different sizes
and shapes must give every grid exactly what the one-grid calls give it
(pools, subcell tables, atlas and images bit for bit), and each grid must
still match the oracle."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _grids():
    from gvdb_amd import workloads as W
    g = np.random.default_rng(77)
    return [
        ("shell", W.sphere_shell(6000, seed=61), 2.0, 0),
        ("slab", W.slab_plane(15000, 1000, seed=62, lo=(40.0, 30.0, 40.0), hi=(160.0, 90.0, 150.0), plane_y=20.0), 2.0, 0),
        ("cloud", W.noise_cloud(5000, seed=63, box=256.0), 1.5, 0),
        ("tiny", (g.random((150, 3)) * 30 + 80).astype(np.float32), 2.0, 0),
        ("thin", W.sphere_shell(3000, seed=64, radius=60.0), 0.5, 0),
        ("orphans", W.sphere_shell(4000, seed=65), 2.0, 2000),   # topology from the first 2000 points
    ]


def _scene(k, w, h):
    from gvdb_amd import workloads as W
    return W.ea_scene(w, h) if k % 2 == 0 else W.iso_scene(w, h, centre=(256.0, 256.0, 256.0), dist=600.0)


def _single(gpu, pts, r, n_topo, k, size):
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(n_topo or len(pts), r, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(4, len(pts), r, H.ZERO3)
    v.GatherDensity(4, len(pts), r, H.ZERO3, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    sc = _scene(k, *size)
    v.apply_scene(sc)
    v.AddRenderBuf(0, size[0], size[1], 16)
    v.Render(sc["shading"], 0, 0)
    img = v.ReadRenderBuf(0, size[0], size[1], 16)
    return v, n, img


SIZES = [(64, 48), (40, 40), (96, 32), (16, 16), (72, 72), (48, 64)]


def test_batch_matches_single_grid_calls(gpu):
    grids = _grids()
    B = gpu.GridBatch(0, 16)
    vols = []
    for name, pts, r, n_topo in grids:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.SetPoints(pts)
        vols.append(v)
    n_all = [len(p) for _, p, _, _ in grids]
    n_topo = [nt or len(p) for _, p, _, nt in grids]
    radii = [r for _, _, r, _ in grids]
    B.RebuildTopology(vols, n_topo, radii)
    B.FinishTopology(vols, False, True)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols, 0)
    lens = B.InsertPointsSubcell(vols, 4, n_all, radii)
    B.GatherDensity(vols, 4, lens, radii, 0)
    B.UpdateApron(vols, 0, 0.0)
    shading = []
    for k, v in enumerate(vols):
        sc = _scene(k, *SIZES[k])
        v.apply_scene(sc)
        v.AddRenderBuf(0, SIZES[k][0], SIZES[k][1], 16)
        shading.append(sc["shading"])
    B.Render(vols, shading, 0, 0)
    B.synchronize()
    assert B.elapsed_ms() > 0
    for k, (name, pts, r, nt) in enumerate(grids):
        s, n, img = _single(gpu, pts, r, nt, k, SIZES[k])
        v = vols[k]
        assert lens[k] == n, name
        H.assert_pools_equal(v, s)
        gs, ss = v.subcells(), s.subcells()
        for key in ("cnt", "off", "pos", "nid"):
            assert np.array_equal(gs[key], ss[key]), (name, key)
        la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
        lb = H.subcell_lists(ss["cnt"], ss["off"], ss["pnt"])
        assert all(np.array_equal(x, y) for x, y in zip(la, lb)), name
        bimg = v.ReadRenderBuf(0, SIZES[k][0], SIZES[k][1], 16)
        if name == "orphans":
            # push-path entries arrive in atomic order: equal only up to ties
            _, rel, zeros = H.density_errors(v.atlas(), s.atlas().astype(np.float64))
            assert zeros == 0 and rel <= 1e-6, name
            assert H.rmse(bimg, img) <= 1e-6, name
        else:
            assert np.array_equal(v.atlas(), s.atlas()), name
            assert np.array_equal(bimg, img), name
        s.close()


def test_batch_grids_match_oracle(gpu):
    import oracle as O
    grids = _grids()[:4]
    B = gpu.GridBatch(0, 8)
    vols = []
    for name, pts, r, _ in grids:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.SetPoints(pts)
        vols.append(v)
    n_all = [len(p) for _, p, _, _ in grids]
    radii = [r for _, _, r, _ in grids]
    B.RebuildTopology(vols, n_all, radii)
    B.FinishTopology(vols)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols)
    lens = B.InsertPointsSubcell(vols, 4, n_all, radii)
    B.GatherDensity(vols, 4, lens, radii)
    B.UpdateApron(vols)
    B.synchronize()
    for k, (name, pts, r, _) in enumerate(grids):
        o, m = H.oracle_volume(O, pts, r, precision=64)
        assert lens[k] == m, name
        H.assert_pools_equal(vols[k], o)
        a, b = vols[k].atlas(), o.atlas(64)
        mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
        _, rel, zeros = H.density_errors(a[mask], b[mask])
        assert zeros == 0 and rel <= 1e-5, (name, rel)


def test_batch_reports_failing_grid(gpu):
    """A stage error names the grid at fault and runs nothing."""
    B = gpu.GridBatch(0, 4)
    good = gpu.VolumeGVDB(0)
    good.Configure(*H.TREE)
    good.AddChannel(0, gpu.T_FLOAT, 1)
    good.SetPoints(np.random.default_rng(3).random((500, 3)).astype(np.float32) * 50 + 100)
    bad = gpu.VolumeGVDB(0)          # not configured
    with pytest.raises(gpu.GVDBError) as e:
        B.RebuildTopology([good, bad], [500, 10], 2.0)
    assert "grid 1" in str(e.value)


def test_batch_gather_variants_match_single_grid_calls(gpu):
    """Radii 2, 2.25, 3 and 5 take the scatter gather's variants (four or five
    voxels per axis of a point's box, any number, 256 threads); two grids of
    each in one batch must leave every atlas exactly as the one-grid calls
    (whose kernels are compiled separately from the multi-grid ones) do."""
    from gvdb_amd import workloads as W
    radii = [2.0, 2.25, 3.0, 5.0, 2.0, 2.25, 3.0, 5.0]
    clouds = [W.sphere_shell(2500 + 500 * k, seed=90 + k) for k in range(len(radii))]
    B = gpu.GridBatch(0, len(radii))
    vols = []
    for pts in clouds:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.SetPoints(pts)
        vols.append(v)
    n_all = [len(p) for p in clouds]
    B.RebuildTopology(vols, n_all, radii)
    B.FinishTopology(vols, False, True)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols, 0)
    lens = B.InsertPointsSubcell(vols, 4, n_all, radii)
    B.GatherDensity(vols, 4, lens, radii, 0)
    B.UpdateApron(vols, 0, 0.0)
    B.synchronize()
    for k, (pts, r) in enumerate(zip(clouds, radii)):
        s = gpu.VolumeGVDB(0)
        s.Configure(*H.TREE)
        s.SetChannelDefault(16, 16, 1)
        s.AddChannel(0, gpu.T_FLOAT, 1)
        s.SetPoints(pts)
        s.RebuildTopology(len(pts), r, H.ZERO3)
        s.FinishTopology(False, True)
        s.UpdateAtlas()
        s.ClearChannel(0)
        n = s.InsertPointsSubcell(4, len(pts), r, H.ZERO3)
        s.GatherDensity(4, len(pts), r, H.ZERO3, n, 0, 0, False)
        s.UpdateApron(0, 0.0)
        assert lens[k] == n, r
        H.assert_pools_equal(vols[k], s)
        a, b = vols[k].atlas(), s.atlas()
        assert a.shape == b.shape and np.array_equal(a, b), "radius %g: batched atlas differs" % r
        s.close()
    for v in vols:
        v.close()
    B.close()


def test_batch_of_130_grids_fp32_and_fp16(gpu):
    """130 tiny grids in one batch (the multi-grid lookup's ballot passes over
    64 grids at a time, so grids 64-129 take its later passes), every third
    one with an FP16 channel, the apron with a nonzero boundary value: each
    grid's pools, atlas and image equal the one-grid calls' bit for bit."""
    g = np.random.default_rng(130)
    G = 130
    grids = []
    for k in range(G):
        n = int(g.integers(40, 400))
        c = g.random(3) * 200 + 40
        grids.append((c + g.normal(0, 6.0, (n, 3))).astype(np.float32))
    types = [gpu.T_HALF if k % 3 == 2 else gpu.T_FLOAT for k in range(G)]
    size = (24, 16)

    def make(pts, t):
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, t, 1)
        v.SetPoints(pts)
        sc = _scene(0, *size)
        v.apply_scene(sc)
        v.AddRenderBuf(0, size[0], size[1], 16)
        return v

    vols = [make(p, t) for p, t in zip(grids, types)]
    ns = [len(p) for p in grids]
    B = gpu.GridBatch(0, G)
    B.RebuildTopology(vols, ns, 2.0)
    B.FinishTopology(vols, False, True)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols, 0)
    lens = B.InsertPointsSubcell(vols, 4, ns, 2.0)
    B.GatherDensity(vols, 4, lens, 2.0, 0)
    B.UpdateApron(vols, 0, 0.25)
    B.Render(vols, gpu.SHADE_VOLUME, 0, 0)
    B.synchronize()
    for k, (pts, t) in enumerate(zip(grids, types)):
        s = make(pts, t)
        s.RebuildTopology(len(pts), 2.0, H.ZERO3)
        s.FinishTopology(False, True)
        s.UpdateAtlas()
        s.ClearChannel(0)
        n = s.InsertPointsSubcell(4, len(pts), 2.0, H.ZERO3)
        s.GatherDensity(4, len(pts), 2.0, H.ZERO3, n, 0, 0, False)
        s.UpdateApron(0, 0.25)
        s.Render(gpu.SHADE_VOLUME, 0, 0)
        assert lens[k] == n, k
        H.assert_pools_equal(vols[k], s)
        dt = np.float16 if t == gpu.T_HALF else np.float32
        a, b = vols[k].atlas(0, dt), s.atlas(0, dt)
        assert a.shape == b.shape and np.array_equal(a, b), "grid %d: batched atlas differs" % k
        assert np.array_equal(vols[k].ReadRenderBuf(0, size[0], size[1], 16), s.ReadRenderBuf(0, size[0], size[1], 16)), k
        s.close()
    for v in vols:
        v.close()
    B.close()
