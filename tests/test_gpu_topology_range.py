"""GPU topology over wide, sparse point sets: the mark paths the build picks
by the size of the points' brick box (a byte per brick up to 2^27 bricks,
direction masks, bits set directly, and above 2^34 bricks no bitmaps at all:
sorted brick keys) give the oracle's pools bit for bit.

The natural cases: a few thousand clusters spread over a 6000^3-voxel box
(a 750^3-brick domain, above 2^27, so bits are set directly), and over a
200000^3-voxel box (1.6e13 bricks: the keyed build), there followed by the
whole splat against the oracle.  Each path forced on a smaller box (the
mark-path knob changes only how the cells are marked), at radius 0 and with
dilation."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _clusters(n_clusters, per, span, seed, sigma=3.0):
    g = np.random.default_rng(seed)
    c = g.uniform(16.0, span - 16.0, (n_clusters, 3))
    p = c[:, None, :] + g.normal(0.0, sigma, (n_clusters, per, 3))
    return np.clip(p.reshape(-1, 3), 0.5, span - 0.5).astype(np.float32)


def _gpu_topology(gpu, pts, radius):
    v, _ = H.gpu_volume(gpu, pts, radius, stop_after="topology")
    return v


@pytest.mark.parametrize("radius", [0.0, 2.0])
def test_sparse_domain_above_2_27_bricks(gpu, radius):
    import oracle as O
    pts = _clusters(3000, 10, 6000.0, 5)
    lo, hi = pts.min(0), pts.max(0)
    bricks = np.prod(np.floor(hi / 8) - np.floor(lo / 8) + 3)
    assert bricks > 2 ** 27
    v = _gpu_topology(gpu, pts, radius)
    o, _ = H.oracle_volume(O, pts, radius, stop_after="topology")
    H.assert_pools_equal(v, o)
    assert v.info()["top_lev"] == o.top_lev()
    v.close()


@pytest.mark.parametrize("radius", [0.0, 2.0, 3.0])
def test_every_mark_path_gives_the_same_pools(gpu, radius):
    import oracle as O
    pts = _clusters(400, 60, 1000.0, 6)
    o, _ = H.oracle_volume(O, pts, radius, stop_after="topology")
    keep = gpu.get_tuning(gpu.KNOB_MARK_PATH)[0]
    try:
        for path in (1, 2, 3, 4, 0):
            gpu.set_tuning(gpu.KNOB_MARK_PATH, path)
            v = _gpu_topology(gpu, pts, radius)
            H.assert_pools_equal(v, o)
            v.close()
    finally:
        gpu.set_tuning(gpu.KNOB_MARK_PATH, keep)


@pytest.mark.parametrize("radius", [0.0, 2.0])
def test_keyed_build_above_2_34_bricks(gpu, radius):
    import oracle as O
    pts = _clusters(2000, 10, 200000.0, 9)
    lo, hi = pts.min(0), pts.max(0)
    assert np.prod(np.floor(hi / 8) - np.floor(lo / 8) + 3) > 2 ** 34
    v, n = H.gpu_volume(gpu, pts, radius)
    o, m = H.oracle_volume(O, pts, radius, precision=64)
    assert o.undefined_events() == 0
    H.assert_pools_equal(v, o)
    assert v.info()["top_lev"] == o.top_lev() == 4
    assert n == m
    a, b = v.atlas(), o.atlas(64)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    _, rel, zeros = H.density_errors(a[mask], b[mask])
    assert zeros == 0 and rel <= 1e-5
    v.close()
