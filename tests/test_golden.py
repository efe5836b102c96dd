"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every stored array (integers exactly, densities and
images to the float32 rounding they were stored with).
GPU: the HIP path, through the C ABI, against the stored arrays: topology,
slots and subcell counts bit-exact; densities within 1e-5 relative of the
stored FP64 values (exact zeros must match); images RMSE <= 1e-3.
"""
import json
import os

import numpy as np
import pytest

import helpers as H

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["sphere_r2", "cloud_r1p5"]
REL_TOL = 1e-5


def _load(name):
    d = dict(np.load(os.path.join(HERE, name + ".npz")))
    meta = json.load(open(os.path.join(HERE, name + ".json")))
    return d, meta


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    import sys
    sys.path.insert(0, HERE)
    import make_golden as G
    d, meta = _load(name)
    out, _ = G.build(name, G.FIXTURES[name])
    assert set(out) == set(d)
    for k, v in out.items():
        v = np.asarray(v)
        if k.startswith("img_"):
            assert np.abs(v - d[k]).max() <= 1e-6, k
        else:
            assert np.array_equal(v, d[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(gpu, name):
    d, meta = _load(name)
    pts, r, scs = d["points"], float(d["radius"]), int(d["scs"])
    v, n = H.gpu_volume(gpu, pts, r, scs=scs)
    assert n == int(d["entries"])
    for lev in range(5):
        p = v.pool(0, lev)
        assert np.array_equal(p["mPos"], d["pos%d" % lev]), "level %d positions" % lev
        assert np.array_equal(p["mParent"], d["parent%d" % lev]), "level %d parents" % lev
        assert np.array_equal(p["mChildList"], d["childref%d" % lev]), "level %d child refs" % lev
        if lev > 0:
            assert np.array_equal(v.pool(1, lev), d["clist%d" % lev]), "level %d child lists" % lev
    leaves = v.pool(0, 0)
    assert np.array_equal(leaves["mValue"], d["slots"])
    assert np.array_equal(v.subcells()["cnt"].astype(np.int64), d["sc_cnt"].astype(np.int64))
    a = v.atlas()
    c = leaves["mValue"] - 1
    got = np.stack([a[z:z + 10, y:y + 10, x:x + 10] for x, y, z in c]).astype(np.float64)
    want = d["bricks"].astype(np.float64)
    dmax, rel, zeros = H.density_errors(got, want)
    assert zeros == 0
    assert rel <= REL_TOL, "max relative density error %.3g" % rel
    for key, shading in (("ea", gpu.SHADE_VOLUME), ("iso", gpu.SHADE_TRILINEAR)):
        sc = meta["scenes"][key]
        v.apply_scene(sc)
        v.AddRenderBuf(0, 64, 64, 16)
        v.Render(shading, 0, 0)
        img = v.ReadRenderBuf(0, 64, 64, 16)
        blk = v.scene_block(64, 64)
        assert np.array_equal(blk[:2], d["scene_" + key][:2])
        assert np.allclose(blk[2:], d["scene_" + key][2:], rtol=1e-5, atol=1e-5)
        e = H.rmse(img, d["img_" + key])
        assert e <= 1e-3, "%s image RMSE %.3g" % (key, e)
    v.close()
