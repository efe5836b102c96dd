"""Point radii wider than a brick: the dilation marks every brick of
[int(p) - r, int(p) + 2r], the insert bins each point into every subcell of
[int(p) - r, int(p) + 1 + r] (the push path: the neighbourhood is wider than
the 27 bricks around the point's own), and the gather takes the general
subcell-list kernel.  Pools, subcell lists and density against the oracle at
radii of 40 and 64 voxels; an entry total past 2^32, where a 32-bit scan of
the subcell counts wraps, is refused before anything is written."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _cloud(n, seed, lo, hi):
    g = np.random.default_rng(seed)
    return (lo + (hi - lo) * g.random((n, 3))).astype(np.float32)


@pytest.mark.parametrize("radius,n", [(40.0, 300), (64.0, 150)])
def test_wide_radius_matches_oracle(gpu, radius, n):
    import oracle as O
    pts = _cloud(n, int(radius), 100.0, 300.0)
    v, m = H.gpu_volume(gpu, pts, radius)
    o, k = H.oracle_volume(O, pts, radius, precision=64)
    assert m == k and m > n * 1000
    H.assert_pools_equal(v, o)
    gs, os_ = v.subcells(), o.subcells(k)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL
    v.close()


def test_entry_total_past_2_32_is_refused(gpu):
    """9000 points in a ball of radius 5 about (256, 256, 256), radius 40,
    one-voxel subcells: each point lands in 82 subcells per axis, all inside
    the dilated topology, 82^3 * 9000 = 4.96e9 entries in all."""
    g = np.random.default_rng(3)
    d = g.standard_normal((9000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = (256.0 + d * 5.0 * g.random((9000, 1)) ** (1 / 3)).astype(np.float32)
    v, _ = H.gpu_volume(gpu, pts, 40.0, stop_after="topology")
    assert v.info()["top_lev"] == 2
    v.UpdateAtlas()
    with pytest.raises(gpu.GVDBError) as e:
        v.InsertPointsSubcell(1, len(pts), 40.0, H.ZERO3)
    assert e.value.code == gpu.ERR_RANGE
    # the grid still inserts what fits
    n = v.InsertPointsSubcell(4, 100, 40.0, H.ZERO3)
    assert n > 0
    v.close()


@pytest.mark.parametrize("radius", [2.0, 12.0])
def test_inserts_reaching_below_the_root_corner(gpu, radius):
    """A topology whose top node is a level-1 node at (192, 192, 192) (points
    about (200, 200, 200), no dilation), and inserted points or radii that
    reach below its corner.  The walk truncates a position less than a cell
    below a node into the node's first cell, so such subcells take a negative
    offset in the brick it finds: the entry lands in a neighbouring subcell,
    or, before the first brick's subcells, nowhere (the oracle counts those as
    undefined).  At radius 2 the points below the corner leave the bucketed
    path for the push path.  Subcell tables equal the oracle's."""
    import oracle as O
    g = np.random.default_rng(8)
    d = g.standard_normal((3000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ball = (200.0 + d * 5.0 * g.random((3000, 1)) ** (1 / 3)).astype(np.float32)
    below = np.stack([g.uniform(186.0, 191.9, 300), g.uniform(186.0, 206.0, 300), g.uniform(193.0, 206.0, 300)], 1)
    pts = np.concatenate([ball, below.astype(np.float32)])
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(len(ball), 0.5, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    n = v.InsertPointsSubcell(4, len(pts), radius, H.ZERO3)
    o = O.Oracle()
    assert o.rebuild_topology(ball, 0.5) == 0 and o.undefined_events() == 0
    assert o.top_lev() == 1
    o.finish_topology()
    o.update_atlas()
    rc, m = o.insert_points_subcell(pts, radius, 4)
    assert rc == 0 and n == m
    assert o.undefined_events() > 0
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    v.close()


def test_insert_past_a_single_leaf_top_node(gpu):
    """A host-built topology of one leaf (every topology point in the brick at
    (200, 200, 200)): the top node is that leaf, and the walk from it returns
    it for every position.  Inserted points around it reach subcells below and
    past the brick; those past its far side have no subcell table entry
    (skipped; the oracle counts them as undefined), and the tables equal the
    oracle's."""
    import oracle as O
    g = np.random.default_rng(12)
    topo = (200.5 + 7.0 * g.random((50, 3))).astype(np.float32)
    extra = (190.0 + 30.0 * g.random((400, 3))).astype(np.float32)
    pts = np.concatenate([topo, extra])
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(topo)
    v.RebuildTopologyCPU(topo)
    v.FinishTopology(True, True)
    assert v.info()["top_lev"] == 0
    v.UpdateAtlas()
    v.SetPoints(pts)
    n = v.InsertPointsSubcell(4, len(pts), 1.0, H.ZERO3)
    o = O.Oracle()
    assert o.rebuild_topology_cpu(topo) == 0
    o.finish_topology()
    assert o.top_lev() == 0
    o.update_atlas()
    rc, m = o.insert_points_subcell(pts, 1.0, 4)
    assert rc == 0 and n == m
    assert o.undefined_events() > 0
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    v.close()
