"""HIP graphs of the per-frame chain ClearChannel -> InsertPointsSubcell ->
GatherDensity -> UpdateApron -> Render, one grid and a batch: a replay gives
the atlas and image of the same calls run eagerly, bit for bit; it re-reads
the point buffer (new positions written in place); it is refused once the
topology changes; a replay whose points leave the topology reports it; calls
outside the chain fail the recording."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

R = 2.0
W = 96


def _grid(gpu, pts, centre):
    from gvdb_amd import workloads as Wk
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(len(pts), R, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.apply_scene(Wk.ea_scene(W, W, centre=centre, dist=300.0))
    v.AddRenderBuf(0, W, W, 16)
    return v


def _chain(gpu, v, n):
    v.ClearChannel(0)
    ln = v.InsertPointsSubcell(4, n, R, H.ZERO3)
    v.GatherDensity(4, n, R, H.ZERO3, ln, 0, 0, False)
    v.UpdateApron(0, 0.0)
    v.Render(gpu.SHADE_VOLUME, 0, 0)
    return ln


def _state(v):
    return v.atlas().view(np.uint32).copy(), v.ReadRenderBuf(0, W, W, 16).view(np.uint32).copy()


def _wipe(gpu, v):
    # overwrite what the chain writes: atlas and render buffer
    v.Compute(gpu.FUNC_FILL_F, 0, 1, (7.0, 0.0, 0.0), False, False, 0.0)
    v.Render(gpu.SHADE_TRILINEAR, 0, 0)


def test_graph_replay_one_grid(gpu):
    from gvdb_amd import workloads as Wk
    c = (128.0, 128.0, 128.0)
    pts = Wk.sphere_shell(40000, seed=41, radius=60.0, centre=c)
    n = len(pts)
    v = _grid(gpu, pts, c)
    ln0 = _chain(gpu, v, n)
    v.synchronize()
    a0, i0 = _state(v)
    assert ln0 > 0 and np.count_nonzero(a0) > 0
    v.CaptureBegin()
    _chain(gpu, v, n)
    g = v.CaptureEnd()
    _wipe(gpu, v)
    assert not np.array_equal(_state(v)[0], a0)
    for _ in range(3):      # replays with eager work between them
        g.launch()
        g.synchronize()
        a1, i1 = _state(v)
        assert np.array_equal(a1, a0) and np.array_equal(i1, i0)
        _wipe(gpu, v)
    # new positions in the same device buffer: the same points in another
    # order give the same atlas (the gather's max does not depend on order)
    rng = np.random.default_rng(5)
    v.SetPoints(pts[rng.permutation(n)])
    g.launch()
    g.synchronize()
    assert np.array_equal(_state(v)[0], a0)
    # other positions inside the topology: the replay equals the eager chain
    sub = pts[:n // 2]
    moved = np.concatenate([sub, sub[::-1] + np.float32(0.25)]).astype(np.float32)
    v.SetPoints(moved)
    g.launch()
    g.synchronize()
    a2, i2 = _state(v)
    assert not np.array_equal(a2, a0)
    ln_e = _chain(gpu, v, n)
    v.synchronize()
    a3, i3 = _state(v)
    assert ln_e > 0 and np.array_equal(a2, a3) and np.array_equal(i2, i3)
    # points outside every brick: the replay reports them
    v.SetPoints(pts + np.float32(400.0))
    g.launch()
    with pytest.raises(gpu.GVDBError) as e:
        g.synchronize()
    assert e.value.code == gpu.ERR_STATE
    # a new topology invalidates the graph
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(n, R, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    with pytest.raises(gpu.GVDBError) as e:
        g.launch()
    assert e.value.code == gpu.ERR_STATE
    g.close()
    v.close()


def test_calls_outside_the_chain_fail_the_recording(gpu):
    from gvdb_amd import workloads as Wk
    c = (64.0, 64.0, 64.0)
    pts = Wk.sphere_shell(5000, seed=42, radius=30.0, centre=c)
    v = _grid(gpu, pts, c)
    _chain(gpu, v, len(pts))
    v.synchronize()
    v.CaptureBegin()
    v.ClearChannel(0)
    with pytest.raises(gpu.GVDBError):
        v.RebuildTopology(len(pts), R, H.ZERO3)
    with pytest.raises(gpu.GVDBError):
        v.CaptureEnd()
    # the grid works eagerly afterwards
    _chain(gpu, v, len(pts))
    v.synchronize()
    v.close()


def test_graph_replay_batch(gpu):
    from gvdb_amd import workloads as Wk
    grids = [Wk.sphere_shell(4000 + 3000 * k, seed=50 + k, radius=20.0 + 8 * k, centre=(80.0 + 20 * k, 90.0, 100.0))
             for k in range(6)]
    c = (130.0, 90.0, 100.0)
    vols = [_grid(gpu, p, c) for p in grids]
    ns = [len(p) for p in grids]
    B = gpu.GridBatch(0, len(vols))

    def chain():
        B.ClearChannel(vols, 0)
        lens = B.InsertPointsSubcell(vols, 4, ns, R)
        B.GatherDensity(vols, 4, lens, R, 0)
        B.UpdateApron(vols, 0, 0.0)
        B.Render(vols, gpu.SHADE_VOLUME, 0, 0)

    chain()
    B.synchronize()
    ref = [_state(v) for v in vols]
    # the batched results equal the one-grid calls'
    for v, p, (a, i) in zip(vols, grids, ref):
        u = _grid(gpu, p, c)
        _chain(gpu, u, len(p))
        u.synchronize()
        au, iu = _state(u)
        assert np.array_equal(au, a) and np.array_equal(iu, i)
        u.close()
    B.CaptureBegin()
    chain()
    g = B.CaptureEnd()
    for v in vols:
        _wipe(gpu, v)
    for _ in range(2):
        g.launch()
        g.synchronize()
        for v, (a, i) in zip(vols, ref):
            a1, i1 = _state(v)
            assert np.array_equal(a1, a) and np.array_equal(i1, i)
        for v in vols:
            _wipe(gpu, v)
    g.close()
    B.close()
    for v in vols:
        v.close()
