"""UpdateApron without allocation (gvdb_update_apron_noalloc) against
gvdb_update_apron, through the C ABI.

Two grids built the same way from the same seeded points: one takes
UpdateApron, the other the allocation-free call; their atlases must be equal
bit for bit.  Two calls in a row on the shared fixed device array, the second
on a smaller grid (fewer bricks, so the array keeps the first call's tail),
each once with the gather's face buffer current and once after an atlas write
(the apron then reads the neighbours' interiors from the atlas).  A grid with
more bricks than the array holds is refused with ERR_OOM and leaves the atlas
as it was.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _splat(g, pts, atlas_type, r=2.0):
    v = g.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, atlas_type, 1)
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(len(pts), r, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(4, len(pts), r, H.ZERO3)
    v.GatherDensity(4, len(pts), r, H.ZERO3, n, 0, 0, False)
    return v


def _pair(gpu, pts, atlas_type, boundval, smooth):
    dt = np.float32 if atlas_type == gpu.T_FLOAT else np.float16
    a, b = _splat(gpu, pts, atlas_type), _splat(gpu, pts, atlas_type)
    if smooth:
        # an atlas write after the gather: the apron reads the atlas, not the face buffer
        for v in (a, b):
            v.Compute(gpu.FUNC_SMOOTH, 0, 1, (0.0, 0.0, 0.0), False, True, 0.0)
    a.UpdateApron(0, boundval)
    b.UpdateApronNoAlloc(0, boundval)
    a.synchronize()
    b.synchronize()
    bricks = len(a.pool(0, 0))
    ref, got = a.atlas(0, dt), b.atlas(0, dt)
    a.close()
    b.close()
    return bricks, ref, got


@pytest.mark.parametrize("smooth", [False, True])
def test_noalloc_apron_matches_update_apron(gpu, smooth):
    from gvdb_amd import workloads as W
    # first call: 60k shell points; second: 4k points (fewer bricks), FP16 channel
    big = W.sphere_shell(60000, seed=11)
    small = W.sphere_shell(4000, seed=12)
    seen = []
    for pts, t, bv in ((big, gpu.T_FLOAT, 0.25), (small, gpu.T_HALF, -0.5)):
        bricks, ref, got = _pair(gpu, pts, t, bv, smooth)
        print("noalloc apron: %d points, %d bricks, type %d, boundval %g, smooth %s"
              % (len(pts), bricks, t, bv, smooth), flush=True)
        assert 0 < bricks <= gpu.NOALLOC_MAX_BRICKS
        assert ref.shape == got.shape
        assert np.array_equal(ref.view(np.uint8), got.view(np.uint8))
        seen.append(bricks)
    assert seen[1] < seen[0]


def test_noalloc_apron_refuses_more_bricks_than_its_array(gpu):
    # one point at the centre of each brick of a 48^3-brick block: 110,592 bricks
    k = np.arange(48, dtype=np.float32) * 8 + 4
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    v = _splat(gpu, pts, gpu.T_FLOAT, r=0.5)
    assert len(v.pool(0, 0)) > gpu.NOALLOC_MAX_BRICKS
    before = v.atlas(0, np.float32)
    with pytest.raises(gpu.GVDBError) as e:
        v.UpdateApronNoAlloc(0, 0.0)
    assert e.value.code == gpu.ERR_OOM
    assert np.array_equal(v.atlas(0, np.float32), before)
    v.UpdateApron(0, 0.0)      # the allocating form still serves it
    v.close()
