"""Seeded random scenes through the whole splat against the oracle: point
clouds of a few shapes and sizes, radii below one voxel (no dilation), near a
brick (the scatter gather's fixed slots) and wider (the general loops),
subcell sizes 2, 4 and 8, and clouds offset from the origin.  Pools and
subcell tables bit-exact, densities within 1e-5 relative of FP64 (all voxels
of the used bricks, apron included)."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _scene(seed):
    g = np.random.default_rng(1000 + seed)
    shape = seed % 3
    n = int(g.integers(2000, 20000))
    c = g.uniform(60.0, 400.0, 3)
    if shape == 0:     # ball
        d = g.standard_normal((n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p = c + d * g.uniform(5.0, 40.0) * g.random((n, 1)) ** (1 / 3)
    elif shape == 1:   # thin sheet
        p = c + np.stack([g.uniform(-40, 40, n), g.normal(0, 0.8, n), g.uniform(-40, 40, n)], 1)
    else:              # a few clusters
        k = int(g.integers(2, 6))
        centres = c + g.uniform(-60, 60, (k, 3))
        p = centres[g.integers(0, k, n)] + g.normal(0, g.uniform(2, 10), (n, 3))
    p = p[(p > 1.0).all(1) & (p < 510.0).all(1)]
    radius = float([0.5, 1.0, 1.5, 2.0, 3.0, 5.5][seed % 6])
    scs = [4, 2, 4, 8][seed % 4]
    return p.astype(np.float32), radius, scs


@pytest.mark.parametrize("seed", range(12))
def test_random_scene_matches_oracle(gpu, seed):
    import oracle as O
    pts, radius, scs = _scene(seed)
    v, n = H.gpu_volume(gpu, pts, radius, scs=scs)
    o, m = H.oracle_volume(O, pts, radius, scs=scs, precision=64)
    assert n == m, (seed, radius, scs, len(pts))
    H.assert_pools_equal(v, o)
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    assert np.array_equal(gs["off"].astype(np.int64), os_["off"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL, (seed, rel, zeros)
    v.close()


@pytest.mark.parametrize("radius,scs", [(1.5, 2), (0.5, 4)])
def test_4cubed_bricks_match_oracle(gpu, radius, scs):
    """Leaf bricks of 4^3 voxels (log2 dims 2, 3, 3, 4, 5): the bucketed
    insert with two subcells per axis or one, the general gather over the
    subcell tables, and the apron's enumeration for a brick width known only
    at run time (6^3 padded bricks); everything against the oracle."""
    import oracle as O
    pts, _, _ = _scene(0)
    dims = (2, 3, 3, 4, 5)
    v = gpu.VolumeGVDB(0)
    v.Configure(*reversed(dims))
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(pts)
    v.RequestFullRebuild(True)
    v.RebuildTopology(len(pts), radius, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(scs, len(pts), radius, H.ZERO3)
    v.GatherDensity(scs, len(pts), radius, H.ZERO3, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    v.synchronize()
    o = O.Oracle(log2dims=dims)
    assert o.rebuild_topology(pts, radius) == 0
    o.finish_topology()
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, m = o.insert_points_subcell(pts, radius, scs)
    assert rc == 0 and n == m
    o.gather_density(radius, scs, 64)
    o.update_apron(0.0, 64)
    H.assert_pools_equal(v, o)
    gs, os_ = v.subcells(), o.subcells(m)
    assert np.array_equal(gs["cnt"].astype(np.int64), os_["cnt"].astype(np.int64))
    la = H.subcell_lists(gs["cnt"], gs["off"], gs["pnt"])
    lb = H.subcell_lists(os_["cnt"], os_["off"], os_["pnt"])
    assert all(np.array_equal(x, y) for x, y in zip(la, lb))
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"], br=6)
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= REL_TOL, (rel, zeros)
    v.close()
