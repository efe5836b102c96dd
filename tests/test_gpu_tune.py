"""The autotuner's launch-shape knobs change time only: every value of every
knob gives the same bits (images, atlases), and each tuner leaves the knob at
one of its values, the fastest it timed.

Knobs: samples fetched ahead (1, 2, 4) and ray-packet shape (8x8, 16x4,
32x2, 64x1) of both marches, threads per brick of the scatter gather (128,
256), bricks per workgroup of the multi-grid apron (1, 2, 4, 8, or the
persistent queue), brick order of the per-brick kernels (launch order,
XCD-contiguous)."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_knobs(gpu):
    keep = [gpu.get_tuning(k)[0] for k in range(gpu.KNOB_COUNT)]
    yield
    for k, v in enumerate(keep):
        gpu.set_tuning(k, v)


def _grid(gpu, pts, t=None):
    v, n = H.gpu_volume(gpu, pts, 2.0, atlas_type=t)
    return v, n


@pytest.mark.parametrize("shading", ["volume", "trilinear"])
def test_render_every_packet_shape_and_unroll_same_pixels(gpu, restore_knobs, shading):
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(30000, seed=21, radius=40.0, centre=(64.0, 64.0, 64.0))
    v, _ = _grid(gpu, pts)
    w, h = 200, 136     # not a multiple of any tile width: partial tiles on both edges
    if shading == "volume":
        sc, mode = W.ea_scene(w, h, centre=(64.0, 64.0, 64.0), dist=200.0), gpu.SHADE_VOLUME
        kt, ku = gpu.KNOB_RENDER_TILE_EA, gpu.KNOB_RENDER_UNROLL_EA
    else:
        sc, mode = W.iso_scene(w, h, centre=(64.0, 64.0, 64.0), dist=200.0), gpu.SHADE_TRILINEAR
        kt, ku = gpu.KNOB_RENDER_TILE_ISO, gpu.KNOB_RENDER_UNROLL_ISO
    v.apply_scene(sc)
    v.AddRenderBuf(0, w, h, 16)
    ref = None
    for tw in (8, 16, 32, 64):
        for u in (1, 2, 4):
            gpu.set_tuning(kt, tw)
            gpu.set_tuning(ku, u)
            v.Render(mode, 0, 0)
            img = v.ReadRenderBuf(0, w, h, 16)
            if ref is None:
                ref = img
                assert np.count_nonzero(np.abs(img - np.asarray(sc["backclr"], np.float32)).max(-1)) > 1000
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (tw, u)
    best, ms = v.TuneRenderShape(mode)
    assert best[0] in (1, 2, 4) and best[1] in (8, 16, 32, 64)
    assert gpu.get_tuning(ku)[0] == best[0] and gpu.get_tuning(kt)[0] == best[1]
    assert len(ms) == 12 and all(t > 0 for t in ms.values())
    assert ms[(best[1], best[0])] == min(ms.values())
    v.Render(mode, 0, 0)
    assert np.array_equal(v.ReadRenderBuf(0, w, h, 16).view(np.uint32), ref.view(np.uint32))
    v.close()


@pytest.mark.parametrize("radius", [2.0, 2.25])
def test_gather_threads_same_atlas(gpu, restore_knobs, radius):
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(40000, seed=22)
    out = {}
    for t in (gpu.T_FLOAT, gpu.T_HALF):
        for nt in (128, 256):
            gpu.set_tuning(gpu.KNOB_GATHER_THREADS, nt)
            gpu.set_tuning(gpu.KNOB_XCD_ORDER, 1 if nt == 256 else 0)
            v, _ = H.gpu_volume(gpu, pts, radius, atlas_type=t)
            out[(t, nt)] = v.atlas(0, np.float32 if t == gpu.T_FLOAT else np.float16)
            if nt == 256:
                assert np.array_equal(out[(t, 128)].view(np.uint8), out[(t, 256)].view(np.uint8)), t
                n = v.InsertPointsSubcell(4, len(pts), radius, H.ZERO3)
                best, ms = v.TuneGather(4, len(pts), radius, H.ZERO3, n, 0)
                assert best in (128, 256) and gpu.get_tuning(gpu.KNOB_GATHER_THREADS)[0] == best
                assert ms[best] == min(ms.values())
                v.UpdateApron(0, 0.0)
                assert np.array_equal(v.atlas(0, out[(t, 128)].dtype).view(np.uint8), out[(t, 128)].view(np.uint8))
            v.close()


def test_one_grid_apron_runs_same_atlas(gpu, restore_knobs):
    """UpdateApron on one grid with runs of 1, 2, 4 and 8 bricks per
    workgroup, launch order and XCD order, float, half and uchar4 channels,
    nonzero boundary value: the same atlas bit for bit."""
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(30000, seed=23, radius=45.0, centre=(90.0, 90.0, 90.0))
    clr = np.random.default_rng(5).integers(0, 1 << 32, len(pts), dtype=np.uint64).astype(np.uint32)
    for t in (gpu.T_FLOAT, gpu.T_HALF):
        ref = None
        for xcd in (0, 1):
            for run in (1, 2, 4, 8):
                gpu.set_tuning(gpu.KNOB_XCD_ORDER, xcd)
                gpu.set_tuning(gpu.KNOB_APRON_RUN_ONE, run)
                v, _ = H.gpu_volume(gpu, pts, 2.0, atlas_type=t, boundval=0.25)
                a = v.atlas(0, np.float32 if t == gpu.T_FLOAT else np.float16).view(np.uint8).copy()
                if ref is None:
                    ref = a
                assert np.array_equal(a, ref), (t, xcd, run)
                v.close()
    # the colour channel's apron (uchar4 words)
    refc = None
    for run in (1, 2, 4, 8):
        gpu.set_tuning(gpu.KNOB_APRON_RUN_ONE, run)
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.AddChannel(1, gpu.T_UCHAR4, 1)
        v.SetPoints(pts, clr)
        v.RequestFullRebuild(True)
        v.RebuildTopology(len(pts), 2.0, H.ZERO3)
        v.FinishTopology(False, True)
        v.UpdateAtlas()
        v.ClearChannel(0)
        v.ClearChannel(1)
        n = v.InsertPointsSubcell(4, len(pts), 2.0, H.ZERO3)
        v.GatherDensity(4, len(pts), 2.0, H.ZERO3, n, 0, 1, False)
        v.UpdateApron(1, 7.0)
        c = v.atlas(1, np.uint32).copy()
        if refc is None:
            refc = c
        assert np.array_equal(c, refc), run
        v.close()


def test_batched_knobs_same_results(gpu, restore_knobs):
    """A 10-grid batch: apron runs of 1, 2, 4, 8 bricks, gather 128/256
    threads and every render shape give the one-grid calls' atlases and
    pixels; the batched tuner picks one of each knob's values."""
    from gvdb_amd import workloads as W
    grids = [W.sphere_shell(3000 + 2000 * k, seed=30 + k, radius=20.0 + 6 * k,
                            centre=(60.0 + 10 * k, 70.0, 80.0)) for k in range(10)]
    w = 96
    ref_atlas, ref_img = [], []
    for pts in grids:
        v, _ = H.gpu_volume(gpu, pts, 2.0)
        v.apply_scene(W.ea_scene(w, w, centre=(100.0, 70.0, 80.0), dist=300.0))
        v.AddRenderBuf(0, w, w, 16)
        v.Render(gpu.SHADE_VOLUME, 0, 0)
        ref_atlas.append(v.atlas())
        ref_img.append(v.ReadRenderBuf(0, w, w, 16))
        v.close()
    vols = []
    for pts in grids:
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.SetPoints(pts)
        v.apply_scene(W.ea_scene(w, w, centre=(100.0, 70.0, 80.0), dist=300.0))
        v.AddRenderBuf(0, w, w, 16)
        vols.append(v)
    ns = [len(p) for p in grids]
    B = gpu.GridBatch(0, len(vols))
    for v in vols:
        v.RequestFullRebuild(True)
    B.RebuildTopology(vols, ns, 2.0)
    B.FinishTopology(vols, False, True)
    B.UpdateAtlas(vols)
    B.ClearChannel(vols, 0)
    lens = B.InsertPointsSubcell(vols, 4, ns, 2.0)
    for run in (1, 2, 4, 8, "queue"):
        for nt in (128, 256):
            gpu.set_tuning(gpu.KNOB_APRON_QUEUE, 1 if run == "queue" else 0)
            gpu.set_tuning(gpu.KNOB_APRON_RUN, 4 if run == "queue" else run)
            gpu.set_tuning(gpu.KNOB_GATHER_THREADS, nt)
            gpu.set_tuning(gpu.KNOB_XCD_ORDER, ((8 if run == "queue" else run) // 2 + nt // 256) & 1)
            B.ClearChannel(vols, 0)
            B.GatherDensity(vols, 4, lens, 2.0, 0)
            B.UpdateApron(vols, 0, 0.0)
            B.synchronize()
            for k, v in enumerate(vols):
                assert np.array_equal(v.atlas().view(np.uint32), ref_atlas[k].view(np.uint32)), (run, nt, k)
    for tw in (8, 16, 32, 64):
        gpu.set_tuning(gpu.KNOB_RENDER_TILE_EA, tw)
        B.Render(vols, gpu.SHADE_VOLUME, 0, 0)
        B.synchronize()
        for k, v in enumerate(vols):
            assert np.array_equal(v.ReadRenderBuf(0, w, w, 16).view(np.uint32), ref_img[k].view(np.uint32)), (tw, k)
    bg, msg = B.Tune(vols, "gather", 4, lens, 2.0)
    ba, msa = B.Tune(vols, "apron")
    br, msr = B.Tune(vols, "render", shading=gpu.SHADE_VOLUME)
    assert bg in (128, 256) and ba in (1, 2, 4, 8, "queue") and br[0] in (1, 2, 4) and br[1] in (8, 16, 32, 64)
    assert gpu.get_tuning(gpu.KNOB_GATHER_THREADS)[0] == bg
    assert gpu.get_tuning(gpu.KNOB_APRON_QUEUE)[0] == (1 if ba == "queue" else 0)
    assert msa[ba] == min(msa.values()) and msg[bg] == min(msg.values())
    B.close()
    for v in vols:
        v.close()
