"""GatherLevelSet on the GPU against the FP64 oracle (gvdbGatherLevelSet,
kernels/cuda_gvdb_particles.cuh:736-774 of the reference): every voxel of a
linked brick holds min(3, min over its subcell's entries of |p - v| - r).

Tolerance: |gpu - fp64| <= 1e-5 |fp64| + 2e-7 on brick interiors (the value
crosses zero at the splat surface, where only an absolute floor is
meaningful).  Cases: the bucket path (8^3 bricks, subcell 4, one and two
subcell rings), the entry-table path (subcell 2), bricks at the origin,
accumulation over a previous level set, and the batched form."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _pts(kind):
    from gvdb_amd import workloads as W
    if kind == "sphere":
        return W.sphere_shell(20000, seed=51)
    if kind == "origin":
        return (np.random.default_rng(52).random((3000, 3)) * 24).astype(np.float32)
    return W.slab_plane(30000, 1500, seed=53, lo=(40.0, 30.0, 40.0), hi=(200.0, 90.0, 160.0), plane_y=20.0)


def _gpu(gpu, pts, r, scs, accumulate_over=None):
    v, n = H.gpu_volume(gpu, pts, r, scs=scs, stop_after="insert")
    if accumulate_over is not None:
        v.GatherLevelSet(scs, len(pts), accumulate_over, H.ZERO3, n, 0, 0, False)
        v.GatherLevelSet(scs, len(pts), r, H.ZERO3, n, 0, 0, True)
    else:
        v.GatherLevelSet(scs, len(pts), r, H.ZERO3, n, 0, 0, False)
    v.synchronize()
    return v, n


def _oracle(pts, r, scs):
    import oracle as O
    o, m = H.oracle_volume(O, pts, r, scs=scs, stop_after="insert")
    o.gather_level_set(r, scs, 64)
    return o, m


def _check(v, o):
    a, b = v.atlas(), o.atlas(64)
    mask = H.interior_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    d = np.abs(a[mask].astype(np.float64) - b[mask])
    tol = 1e-5 * np.abs(b[mask]) + 2e-7
    bad = d > tol
    assert not bad.any(), "%d voxels off, worst %.3g (value %.3g)" % (bad.sum(), d.max(), b[mask][np.argmax(d)])
    assert (a[mask] <= 3.0).all() and (a[mask] < 0).any()


@pytest.mark.parametrize("kind,r,scs", [("sphere", 2.0, 4), ("slab", 2.0, 4), ("sphere", 5.0, 4),
                                        ("sphere", 2.0, 2), ("origin", 2.0, 4)])
def test_level_set_vs_fp64(gpu, kind, r, scs):
    pts = _pts(kind)
    v, n = _gpu(gpu, pts, r, scs)
    o, m = _oracle(pts, r, scs)
    assert n == m
    _check(v, o)
    v.close()


def test_level_set_accumulates(gpu):
    """A second gather with bAccum starts from the first one's values:
    min(first, second), which for a smaller radius second is the first."""
    import oracle as O
    pts = _pts("sphere")
    v, n = _gpu(gpu, pts, 2.0, 4, accumulate_over=1.0)
    o, m = H.oracle_volume(O, pts, 2.0, stop_after="insert")
    o.gather_level_set(1.0, 4, 64)
    o.gather_level_set(2.0, 4, 64, accumulate=True)
    _check(v, o)
    v.close()


def test_level_set_batched_matches_single(gpu):
    """Two grids through GridBatch give the single-grid values bit for bit."""
    pts = [_pts("sphere"), _pts("slab")]
    singles = [_gpu(gpu, p, 2.0, 4)[0] for p in pts]
    vols, lens = [], []
    for p in pts:
        v, n = H.gpu_volume(gpu, p, 2.0, stop_after="insert")
        vols.append(v)
        lens.append(n)
    B = gpu.GridBatch(0, 2)
    B.GatherLevelSet(vols, 4, lens, 2.0, 0)
    B.synchronize()
    for v, s in zip(vols, singles):
        assert np.array_equal(v.atlas(), s.atlas())
        v.close()
        s.close()
    B.close()


@pytest.mark.parametrize("rg", [1.5, 3.0])
def test_density_at_other_radius_after_pull_insert(gpu, rg):
    """GatherDensity at a radius other than the insert's reads the entry
    tables, which a bucketed (8^3 brick, subcell 4) insert builds only on
    demand; the gather must build them before it reads them.  At a larger
    radius than the insert's, every entry of a subcell can raise a voxel."""
    import oracle as O
    pts = _pts("sphere")
    v, n = H.gpu_volume(gpu, pts, 2.0, stop_after="insert")
    v.GatherDensity(4, len(pts), rg, H.ZERO3, n, 0, 0, False)
    v.synchronize()
    o, m = H.oracle_volume(O, pts, 2.0, stop_after="insert")
    o.gather_density(rg, 4, 64)
    assert n == m
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    dmax, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= 1e-5, "max relative density error %.3g (abs %.3g)" % (rel, dmax)
    v.close()


@pytest.mark.parametrize("scs,r", [(4, 2.0), (4, 1.0), (2, 2.0)])
def test_level_set_empty_lists(gpu, scs, r):
    """An insert of no points leaves every subcell empty; the level set is
    then 3 on every voxel of every linked brick, as in the reference, which
    launches its gather without checking the entry count.  Cases: the
    bucket path, the bucket insert read through the entry tables (gather
    radius 1 after an insert at 2), and the unbucketed insert (subcell 2)."""
    import oracle as O
    pts = _pts("sphere")
    v, _ = H.gpu_volume(gpu, pts, 2.0, scs=scs, stop_after="topology")
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(scs, 0, 2.0, H.ZERO3)
    assert n == 0
    v.GatherLevelSet(scs, 0, r, H.ZERO3, 0, 0, 0, False)
    v.synchronize()
    o, _ = H.oracle_volume(O, pts, 2.0, stop_after="topology")
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, m = o.insert_points_subcell(pts[:0], 2.0, scs)
    assert rc == 0 and m == 0
    o.gather_level_set(r, scs, 64)
    a, b = v.atlas(), o.atlas(64)
    mask = H.interior_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    assert (b[mask] == 3.0).all()
    assert np.array_equal(a[mask], b[mask].astype(np.float32))
    v.close()


# ------------------------------------------------ FP16 subcell variants
def _translated():
    """Points q with w = q + t; 50 have q.x < 0 <= w.x (binned by the FP16
    rule only) and 10 have q.z = NOHIT; the topology is built from w."""
    base = _pts("sphere")
    t = np.array([5.0, 3.0, 0.0], np.float32)
    q = (base - t).astype(np.float32)
    q[:50, 0] = np.float32(-0.5)
    q[50:60, 2] = np.float32(1.0e10)
    w = (q + t).astype(np.float32)
    topo = np.concatenate([base, w[:50]]).astype(np.float32)
    return topo, q, t


def _gpu16(gpu, topo, q, t, r, scs):
    v, _ = H.gpu_volume(gpu, topo, r, scs=scs, stop_after="topology")
    v.UpdateAtlas()
    v.ClearChannel(0)
    v.SetPoints(q)
    n = v.InsertPointsSubcell_FP16(scs, len(q), r, t)
    return v, n


def _oracle16(topo, q, t, r, scs):
    import oracle as O
    o, _ = H.oracle_volume(O, topo, r, stop_after="topology")
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, m = o.insert_points_subcell_fp16(q, r, scs, t)
    assert rc == 0
    return o, m


FP16_CASES = {"sphere_s4": 4, "sphere_s2": 2, "translated_s4": 4}


def _case16(name):
    if name.startswith("translated"):
        return _translated()
    p = _pts("sphere")
    return p, p, np.zeros(3, np.float32)


@pytest.mark.parametrize("case", sorted(FP16_CASES))
def test_fp16_insert_bit_exact(gpu, case):
    """Counts, offsets and the u16 entries (as a sorted set per subcell) of
    InsertPointsSubcell_FP16 equal the oracle's."""
    scs = FP16_CASES[case]
    topo, q, t = _case16(case)
    v, n = _gpu16(gpu, topo, q, t, 2.0, scs)
    o, m = _oracle16(topo, q, t, 2.0, scs)
    assert n == m
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], v.subcells_fp16())
    lb = H.subcell_lists(os_["cnt"], os_["off"], o.subcells_fp16(m))
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    v.close()


@pytest.mark.parametrize("case", sorted(FP16_CASES))
def test_fp16_level_set_vs_fp64(gpu, case):
    """GatherLevelSet_FP16 within 1e-5 relative (2e-7 absolute at the zero
    crossing) of the FP64 level set over the same dequantised entries."""
    scs = FP16_CASES[case]
    topo, q, t = _case16(case)
    v, n = _gpu16(gpu, topo, q, t, 2.0, scs)
    v.GatherLevelSet_FP16(scs, len(q), 2.0, t, n, 0, 0)
    v.synchronize()
    o, m = _oracle16(topo, q, t, 2.0, scs)
    assert o.gather_level_set_fp16(2.0, scs, 64) == 0
    _check(v, o)
    v.close()


def test_fp16_gather_needs_the_fp16_insert(gpu):
    """The u16 entries belong to the insert that made them: a later plain
    insert makes GatherLevelSet_FP16 an error until InsertPointsSubcell_FP16
    runs again."""
    from gvdb_amd import GVDBError
    pts = _pts("sphere")
    v, n = H.gpu_volume(gpu, pts, 2.0, stop_after="insert")
    with pytest.raises(GVDBError):
        v.GatherLevelSet_FP16(4, len(pts), 2.0, H.ZERO3, n, 0, 0)
    n16 = v.InsertPointsSubcell_FP16(4, len(pts), 2.0, H.ZERO3)
    assert n16 == n
    v.GatherLevelSet_FP16(4, len(pts), 2.0, H.ZERO3, n16, 0, 0)
    v.InsertPointsSubcell(4, len(pts), 2.0, H.ZERO3)
    with pytest.raises(GVDBError):
        v.GatherLevelSet_FP16(4, len(pts), 2.0, H.ZERO3, n, 0, 0)
    v.close()
