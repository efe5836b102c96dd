"""The point path's colour channel against the oracle: SetPoints' colour
buffer carried by InsertPointsSubcell, GatherDensity / GatherLevelSet writing
the uchar4 channel chanClr (each subcell's mean colour), UpdateApron on the
uchar4 channel, and the march reading it through SetColorChannel.  The call
sequences are the reference's colour samples' (gSprayDeposit: density 0,
colour 1, accumulating gathers; gFluidSurface: level set 0, colour 1,
UpdateApron(1, 0)).

Bars: the colour channel bit for bit against the oracle's exact mean (the
oracle pins that mean against the reference's float formula to one unit);
the density channel unchanged by the colours; images within 1e-3 RMSE of the
oracle's FP32 march over the same atlases."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

R = 2.0


def _colors(n, seed):
    return np.random.default_rng(seed).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


def _gpu(gpu, pts, clr, levelset=False):
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.AddChannel(1, gpu.T_UCHAR4, 1)
    v.SetPoints(pts, clr)
    v.RequestFullRebuild(True)
    v.RebuildTopology(len(pts), R, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    v.ClearChannel(1)
    return v


def _oracle(O, pts, clr):
    o = O.Oracle()
    assert o.rebuild_topology(pts, R) == 0
    o.finish_topology()
    assert o.update_atlas() == 0
    o.add_color_channel()
    o.clear_channel()
    return o


@pytest.mark.parametrize("seed", [1, 2])
def test_spray_deposit_sequence(gpu, seed):
    """Density 0 and colour 1, a plain gather then an accumulating one (as the
    spray sample deposits each frame), UpdateApron on both channels."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(20000, seed=70 + seed, radius=50.0, centre=(96.0, 96.0, 96.0))
    clr = _colors(len(pts), seed)
    v = _gpu(gpu, pts, clr)
    o = _oracle(O, pts, clr)
    for accumulate in (False, True):
        n = v.InsertPointsSubcell(4, len(pts), R, H.ZERO3)
        rc, m = o.insert_points_subcell(pts, R, 4, colors=clr)
        assert rc == 0 and n == m
        v.GatherDensity(4, len(pts), R, H.ZERO3, n, 0, 1, accumulate)
        o.gather_density(R, 4, 64, accumulate)
        o.gather_color(4, accumulate, exact=True)
        v.UpdateApron(0, 0.0)
        v.UpdateApron(1, 0.0)
        o.update_apron(0.0, 64)
        o.update_apron_color(0.0)
        got, want = v.atlas(1, np.uint32), o.color_atlas()
        assert np.count_nonzero(want) > 1000
        assert np.array_equal(got, want), "accumulate=%s: %d voxels differ" % (accumulate, np.count_nonzero(got != want))
        info = o.atlas_info()
        _, rel, zeros = H.brick_density_errors(v.atlas(), o.atlas(64), o.atlas_map(), info["cnt"])
        assert zeros == 0 and rel <= 1e-5
    v.close()


def test_colors_leave_the_density_unchanged_and_boundary_bytes(gpu):
    from gvdb_amd import workloads as W
    import oracle as O
    pts = W.sphere_shell(15000, seed=81, radius=40.0, centre=(80.0, 80.0, 80.0))
    clr = _colors(len(pts), 9)
    a, _ = H.gpu_volume(gpu, pts, R)
    v = _gpu(gpu, pts, clr)
    n = v.InsertPointsSubcell(4, len(pts), R, H.ZERO3)
    v.GatherDensity(4, len(pts), R, H.ZERO3, n, 0, 1, False)
    v.UpdateApron(0, 0.0)
    assert np.array_equal(v.atlas().view(np.uint32), a.atlas().view(np.uint32))
    o = _oracle(O, pts, clr)
    o.insert_points_subcell(pts, R, 4, colors=clr)
    o.gather_color(4, False, exact=True)
    for bv in (3.7, -2.0, 300.0):
        v.UpdateApron(1, bv)
        o.update_apron_color(bv)
        assert np.array_equal(v.atlas(1, np.uint32), o.color_atlas()), bv
    # and the allocation-free apron on the colour channel
    v.UpdateApronNoAlloc(1, 5.0)
    o.update_apron_color(5.0)
    assert np.array_equal(v.atlas(1, np.uint32), o.color_atlas())
    a.close()
    v.close()


def test_fluid_surface_sequence_level_set_with_color(gpu):
    """gFluidSurface: GatherLevelSet into 0 with colour 1, UpdateApron(0, 3),
    UpdateApron(1, 0)."""
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(12000, seed=91, radius=35.0, centre=(70.0, 70.0, 70.0))
    clr = _colors(len(pts), 11)
    v = _gpu(gpu, pts, clr)
    o = _oracle(O, pts, clr)
    n = v.InsertPointsSubcell(4, len(pts), R, H.ZERO3)
    o.insert_points_subcell(pts, R, 4, colors=clr)
    v.GatherLevelSet(4, len(pts), R, H.ZERO3, n, 0, 1, False)
    o.gather_level_set(R, 4, 64)
    o.gather_color(4, False, exact=True)
    v.UpdateApron(0, 3.0)
    v.UpdateApron(1, 0.0)
    o.update_apron(3.0, 64)
    o.update_apron_color(0.0)
    assert np.array_equal(v.atlas(1, np.uint32), o.color_atlas())
    v.close()


@pytest.mark.parametrize("shading", ["volume", "trilinear"])
def test_render_with_color_channel(gpu, shading):
    import oracle as O
    from gvdb_amd import workloads as W
    c = (96.0, 96.0, 96.0)
    pts = W.sphere_shell(25000, seed=95, radius=50.0, centre=c)
    clr = _colors(len(pts), 13)
    v = _gpu(gpu, pts, clr)
    n = v.InsertPointsSubcell(4, len(pts), R, H.ZERO3)
    v.GatherDensity(4, len(pts), R, H.ZERO3, n, 0, 1, False)
    v.UpdateApron(0, 0.0)
    v.UpdateApron(1, 0.0)
    v.SetColorChannel(1)
    w = 160
    if shading == "volume":
        sc, mode = W.ea_scene(w, w, centre=c, dist=320.0), gpu.SHADE_VOLUME
    else:
        sc, mode = W.iso_scene(w, w, centre=c, dist=320.0), gpu.SHADE_TRILINEAR
    v.apply_scene(sc)
    v.AddRenderBuf(0, w, w, 16)
    v.Render(mode, 0, 0)
    img = v.ReadRenderBuf(0, w, w, 16)
    v.SetColorChannel(-1)
    v.Render(mode, 0, 0)
    plain = v.ReadRenderBuf(0, w, w, 16)
    assert not np.array_equal(img, plain)
    o = _oracle(O, pts, clr)
    o.set_atlas(v.atlas(), 32)
    o.set_color_atlas(v.atlas(1, np.uint32))
    o.set_color_channel(True)
    ref = o.render(v.scene_block(w, w), O.linear_tf(*sc["tf"]), 7 if shading == "volume" else 4, 32)
    e = H.rmse(img, ref)
    assert e <= 1e-3, "RMSE %.3g" % e
    v.close()
