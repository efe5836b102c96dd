"""The serving loop: the same timed request mix served one at a time (plain)
and with continuous batching must give every request the same result --
each render's pixels bit for bit -- and leave every grid in the same state
as building it from scratch from all its points."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _make(gpu, budget=200 << 30):
    from gvdb_amd import workloads as W
    shapes = [W.sphere_shell(8000, seed=81), W.slab_plane(9000, 1000, seed=82, lo=(40.0, 30.0, 40.0),
                                                          hi=(140.0, 80.0, 120.0), plane_y=20.0),
              W.noise_cloud(6000, seed=83, box=256.0), W.sphere_shell(3000, seed=84, radius=50.0),
              W.noise_cloud(5000, seed=85, box=200.0)]
    S = gpu.Server(0, budget, 16)
    vols = []
    for k, pts in enumerate(shapes):
        v = gpu.VolumeGVDB(0)
        v.Configure(*H.TREE)
        v.SetChannelDefault(16, 16, 1)
        v.AddChannel(0, gpu.T_FLOAT, 1)
        v.apply_scene(W.ea_scene(64, 64))
        assert S.add_grid(v, pts, 2.0) == k
        vols.append(v)
    S.reset()
    return S, vols, shapes


def _requests(S, shapes, n=36, seed=9, span=0.05):
    g = np.random.default_rng(seed)
    arrivals = np.sort(g.random(n) * span)
    inserts = {}
    for t in arrivals:
        k = int(g.integers(0, len(shapes)))
        if g.random() < 0.7:
            w = int(g.integers(2, 21)) * 8
            h = int(g.integers(2, 21)) * 8
            S.submit_render(k, t, w, h, angs=(float(g.uniform(-180, 180)), float(g.uniform(-60, 60)), 0.0))
        else:
            base = shapes[k][g.integers(0, len(shapes[k]), int(g.integers(300, 2000)))]
            pts = (base + g.normal(0, 2.0, base.shape)).astype(np.float32)
            pts = np.maximum(pts, 1.0).astype(np.float32)
            S.submit_insert(k, t, pts)
            inserts.setdefault(k, []).append(pts)
    return inserts


def test_served_results_equal_alone_and_in_the_mix(gpu):
    S, vols, shapes = _make(gpu)
    inserts = _requests(S, shapes)
    lat0, st0, wall0 = S.run(S.PLAIN)
    assert np.all(st0 == 0)
    plain = [r[2].copy() for r in S.requests if r[0] == "render"]
    S.reset()
    lat1, st1, wall1 = S.run(S.CONTINUOUS)
    assert np.all(st1 == 0)
    mixed = [r[2] for r in S.requests if r[0] == "render"]
    assert len(plain) == len(mixed) and len(plain) > 10
    for a, b in zip(plain, mixed):
        assert np.array_equal(a, b)
    assert np.all(lat0 >= 0) and np.all(lat1 >= 0)
    # each grid ends as if built alone from all its points
    for k, v in enumerate(vols):
        pts = np.concatenate([shapes[k]] + inserts.get(k, []))
        ref, n = H.gpu_volume(gpu, pts, 2.0)
        H.assert_pools_equal(v, ref)
        assert np.array_equal(v.atlas(), ref.atlas())
        ref.close()


def test_insert_over_budget_is_refused(gpu):
    S, vols, shapes = _make(gpu)
    used = S.device_bytes()
    S.close()
    S, vols, shapes = _make(gpu, budget=used + (1 << 20))
    big = np.repeat(shapes[0], 40, axis=0)          # far beyond the 1 MiB headroom
    S.submit_render(0, 0.0, 64, 64)
    S.submit_insert(1, 0.0, big)
    S.submit_render(2, 0.0, 32, 32)
    lat, st, wall = S.run(S.CONTINUOUS)
    assert st[0] == 0 and st[2] == 0
    assert st[1] == gpu.ERR_OOM
