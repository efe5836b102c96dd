"""Incremental topology on the GPU against the oracle (ActivateIncreBricksGPU,
src/gvdb_volume_gvdb.cpp:1020-1156 of the reference; AccumulateTopology
:1158-1178): pools and child lists bit-exact after RebuildTopology's
incremental branch (the third call after Configure) and after
AccumulateTopology, the no-new-node case, and the gather over an incremental
tree within 1e-5 relative of the FP64 oracle."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _ball(n, seed, centre=(200.0, 180.0, 220.0), radius=30.0):
    g = np.random.default_rng(seed)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius + g.standard_normal(n)
    return (np.asarray(centre) + d * r[:, None]).astype(np.float32)


def _pair(shift, n=20000):
    A = _ball(n, 61)
    B = _ball(n, 62, centre=(200.0 + shift[0], 180.0 + shift[1], 220.0 + shift[2]))
    return A, B


def _gpu_steps(gpu, A, B, r, accum):
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(A)
    v.RebuildTopology(len(A), r, H.ZERO3)
    if not accum:
        v.RebuildTopology(len(A), r, H.ZERO3)
    v.SetPoints(B)
    if accum:
        v.AccumulateTopology(len(B), r, H.ZERO3)
    else:
        v.RebuildTopology(len(B), r, H.ZERO3)
    return v


def _oracle_steps(A, B, r, accum):
    import oracle as O
    o = O.Oracle()
    assert o.rebuild_topology(A, r) == 0
    if not accum:
        assert o.rebuild_topology(A, r) == 0
    rc = o.accumulate_topology(B, r) if accum else o.rebuild_topology(B, r)
    assert rc == 0
    return o


CASES = [((15.0, 0.0, 0.0), False), ((70.0, 0.0, 0.0), False), ((15.0, 0.0, 0.0), True),
         ((70.0, -40.0, 0.0), True)]


@pytest.mark.parametrize("shift,accum", CASES)
def test_incremental_pools_bit_exact(gpu, shift, accum):
    A, B = _pair(shift)
    v = _gpu_steps(gpu, A, B, 2.0, accum)
    o = _oracle_steps(A, B, 2.0, accum)
    assert o.undefined_events() == 0
    assert len(o.pool(0, 0)) > len(_oracle_full(A, 2.0).pool(0, 0))
    H.assert_pools_equal(v, o)
    assert v.info()["top_lev"] == o.top_lev()
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    o.finish_topology()
    assert o.update_atlas() == 0
    assert v.atlas_info() == o.atlas_info()        # used = flagged leaves
    H.assert_pools_equal(v, o)                     # mValue of the flagged leaves
    v.close()


def _oracle_full(pts, r):
    import oracle as O
    o = O.Oracle()
    assert o.rebuild_topology(pts, r) == 0
    return o


def test_incremental_without_new_nodes_changes_nothing(gpu):
    A, _ = _pair((0.0, 0.0, 0.0))
    v = _gpu_steps(gpu, A, A[:500].copy(), 2.0, False)
    o = _oracle_steps(A, A[:500].copy(), 2.0, False)
    f = _oracle_full(A, 2.0)
    H.assert_pools_equal(v, o)
    H.assert_pools_equal(v, f)
    v.close()


@pytest.mark.parametrize("shift,accum", [((70.0, 0.0, 0.0), False), ((70.0, -40.0, 0.0), True)])
def test_gather_over_incremental_tree_vs_fp64(gpu, shift, accum):
    """InsertPointsSubcell + GatherDensity + UpdateApron on the incremental
    tree: unflagged (delinked) leaves get no subcells and no atlas slot."""
    A, B = _pair(shift)
    v = _gpu_steps(gpu, A, B, 2.0, accum)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(4, len(B), 2.0, H.ZERO3)
    v.GatherDensity(4, len(B), 2.0, H.ZERO3, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    v.synchronize()
    o = _oracle_steps(A, B, 2.0, accum)
    o.finish_topology()
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, m = o.insert_points_subcell(B, 2.0, 4)
    assert rc == 0 and n == m
    o.gather_density(2.0, 4, 64)
    o.update_apron(0.0, 64)
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    dmax, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= 1e-5, "max relative density error %.3g (abs %.3g)" % (rel, dmax)
    v.close()


def test_accumulate_refuses_a_deeper_tree(gpu):
    """Points needing a deeper tree than the current root: the reference
    would walk past its root; here AccumulateTopology reports GVDB_ERR_DEPTH."""
    from gvdb_amd import GVDBError
    A, _ = _pair((0.0, 0.0, 0.0), n=2000)
    far = np.concatenate([A, A + np.float32(600.0)]).astype(np.float32)     # spans two level-2 nodes
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(A)
    v.RebuildTopology(len(A), 2.0, H.ZERO3)
    v.SetPoints(far)
    with pytest.raises(GVDBError):
        v.AccumulateTopology(len(far), 2.0, H.ZERO3)
    v.close()


@pytest.mark.parametrize("shift,accum", [((70.0, 0.0, 0.0), False), ((70.0, -40.0, 0.0), True)])
def test_incremental_keyed_path_same_pools(gpu, shift, accum):
    """The keyed form of the incremental build (missing cells as sorted keys,
    taken above 2^34 cells per level; forced here with the mark-path knob)
    gives the oracle's pools."""
    A, B = _pair(shift)
    keep = gpu.get_tuning(gpu.KNOB_MARK_PATH)[0]
    try:
        gpu.set_tuning(gpu.KNOB_MARK_PATH, 4)
        v = _gpu_steps(gpu, A, B, 2.0, accum)
    finally:
        gpu.set_tuning(gpu.KNOB_MARK_PATH, keep)
    o = _oracle_steps(A, B, 2.0, accum)
    assert o.undefined_events() == 0
    H.assert_pools_equal(v, o)
    assert v.info()["top_lev"] == o.top_lev()
    v.close()


def _clusters(n_clusters, per, span, seed, sigma=3.0):
    g = np.random.default_rng(seed)
    c = g.uniform(16.0, span - 16.0, (n_clusters, 3))
    p = c[:, None, :] + g.normal(0.0, sigma, (n_clusters, per, 3))
    return np.clip(p.reshape(-1, 3), 0.5, span - 0.5).astype(np.float32)


@pytest.mark.parametrize("accum", [False, True])
def test_incremental_above_2_34_cells(gpu, accum):
    """Clusters over a 200000^3-voxel box (1.6e13 bricks), then more clusters
    in the same box: the incremental branch's boxes exceed 2^34 cells, so the
    keyed form runs; pools equal the oracle's."""
    A = _clusters(2000, 10, 200000.0, 9)
    B = _clusters(1500, 10, 200000.0, 19)
    v = _gpu_steps(gpu, A, B, 2.0, accum)
    o = _oracle_steps(A, B, 2.0, accum)
    assert o.undefined_events() == 0
    assert len(o.pool(0, 0)) > len(_oracle_full(A, 2.0).pool(0, 0))
    H.assert_pools_equal(v, o)
    assert v.info()["top_lev"] == o.top_lev()
    v.close()
