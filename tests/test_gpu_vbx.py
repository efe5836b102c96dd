"""SaveVBX / LoadVBX on the GPU (src/gvdb_volume_gvdb.cpp:1626-1768 and
:507-685 of the reference).  The file SaveVBX writes equals, byte for byte,
the file the format restatement (oracle/vbx.py) writes from the same pools
and atlas; LoadVBX restores pools, atlas and rendered pixels exactly, also for
a tree with leaves the incremental build deactivated; malformed files fail
with GVDB_ERR_IO and bitmask files with GVDB_ERR_UNSUPPORTED; a file's
grid transform renders as SetTransform's."""
import struct

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _sphere():
    from gvdb_amd import workloads as W
    return W.sphere_shell(20000, seed=81)


def _meta(v):
    import model as M
    res, rng = M.tree_ranges((3, 3, 3, 4, 5))
    cnt = [len(v.pool(0, l)) for l in range(5)]
    top = max(l for l in range(5) if cnt[l] > 0)
    return dict(levels=5, root=M.elem(0, top, 0), log2dims=(3, 3, 3, 4, 5), res=res, ranges=rng, apron=1, brick=8,
                axiscnt=v.atlas_info()["cnt"], bricks=cnt[0])


def _expected_bytes(v):
    import vbx
    p0 = [v.pool(0, l).tobytes() for l in range(5)]
    p1 = [v.pool(1, l).tobytes() if l else b"" for l in range(5)]
    return vbx.write_vbx(_meta(v), p0, p1, [(3, v.atlas())])


def _render(v):
    from gvdb_amd import workloads as W
    sc = W.iso_scene(96, 80, centre=(256.0, 256.0, 256.0), dist=600.0)
    v.apply_scene(sc)
    v.AddRenderBuf(0, sc["width"], sc["height"], 16)
    v.Render(4, 0, 0)
    return v.ReadRenderBuf(0, sc["width"], sc["height"], 16)


def test_save_vbx_bytes_match_the_format(gpu, tmp_path):
    v, _ = H.gpu_volume(gpu, _sphere(), 2.0)
    f = tmp_path / "sphere.vbx"
    v.SaveVBX(str(f))
    got = f.read_bytes()
    assert got == _expected_bytes(v)
    v.close()


def test_load_vbx_round_trip(gpu, tmp_path):
    v, _ = H.gpu_volume(gpu, _sphere(), 2.0)
    f = tmp_path / "sphere.vbx"
    v.SaveVBX(str(f))
    img = _render(v)
    w = gpu.VolumeGVDB(0)
    w.LoadVBX(str(f))
    H.assert_pools_equal(w, v)
    assert w.atlas_info() == v.atlas_info()
    assert np.array_equal(w.atlas(), v.atlas())
    assert np.array_equal(w.atlas_map().view(np.int32), v.atlas_map().view(np.int32))
    assert np.array_equal(_render(w), img)
    assert np.array_equal(w.transform(), np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1], [0, 0, 0]], np.float32))
    v.close()
    w.close()


def test_load_vbx_of_an_incremental_tree(gpu, tmp_path):
    """Unflagged leaves travel in the pools; after loading, UpdateAtlas gives
    the flagged ones the same slots, so atlas and pixels match.  LoadVBX
    counts every stored leaf as used (PoolRead), so UpdateAtlas grows the
    atlas in Z to hold them all, as the reference's AtlasResize does; the
    added slabs stay empty."""
    from gvdb_amd import workloads as W
    A = W.sphere_shell(20000, seed=82)
    B = (A + np.float32([70.0, 0.0, 0.0])).astype(np.float32)
    v = gpu.VolumeGVDB(0)
    v.Configure(*H.TREE)
    v.SetChannelDefault(16, 16, 1)
    v.AddChannel(0, gpu.T_FLOAT, 1)
    v.SetPoints(A)
    v.RebuildTopology(len(A), 2.0, H.ZERO3)
    v.RebuildTopology(len(A), 2.0, H.ZERO3)
    v.SetPoints(B)
    v.RebuildTopology(len(B), 2.0, H.ZERO3)
    v.FinishTopology(False, True)
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(4, len(B), 2.0, H.ZERO3)
    v.GatherDensity(4, len(B), 2.0, H.ZERO3, n, 0, 0, False)
    v.UpdateApron(0, 0.0)
    leaves = v.pool(0, 0)
    assert (leaves["mFlags"] == 0).any()
    f = tmp_path / "incr.vbx"
    v.SaveVBX(str(f))
    assert f.read_bytes() == _expected_bytes(v)
    img = _render(v)
    w = gpu.VolumeGVDB(0)
    w.LoadVBX(str(f))
    for lev in range(5):
        a, b = w.pool(0, lev), v.pool(0, lev)
        for fld in ("mLev", "mFlags", "mPos", "mParent", "mChildList"):
            assert np.array_equal(a[fld], b[fld])
        fl = b["mFlags"] == 1
        assert np.array_equal(a["mValue"][fl], b["mValue"][fl])
    aw, av = w.atlas(), v.atlas()
    z = av.shape[0]
    assert aw.shape[0] > z and aw.shape[1:] == av.shape[1:]
    assert np.array_equal(aw[:z], av) and not aw[z:].any()
    assert np.array_equal(_render(w), img)
    v.close()
    w.close()


def test_load_vbx_errors(gpu, tmp_path):
    from gvdb_amd import GVDBError
    import vbx
    v, _ = H.gpu_volume(gpu, _sphere(), 2.0)
    f = tmp_path / "ok.vbx"
    v.SaveVBX(str(f))
    good = f.read_bytes()
    w = gpu.VolumeGVDB(0)
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(str(tmp_path / "missing.vbx"))
    assert e.value.code == -9
    (tmp_path / "short.vbx").write_bytes(good[: len(good) // 2])
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(str(tmp_path / "short.vbx"))
    assert e.value.code == -9
    # version 2.0 with the bitmask flag set
    hdr = bytearray(good[:54])
    hdr[0], hdr[1] = 2, 0
    masked = bytes(hdr) + b"\x01" + good[54:]
    (tmp_path / "masked.vbx").write_bytes(masked)
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(str(tmp_path / "masked.vbx"))
    assert e.value.code == -7
    # a file's non-identity transform is the grid's: it renders as the same
    # grid given that transform with SetTransform
    xf = bytearray(good)
    struct.pack_into("<3f", xf, 2 + 24, 2.0, 2.0, 2.0)
    (tmp_path / "scaled.vbx").write_bytes(bytes(xf))
    w.LoadVBX(str(tmp_path / "scaled.vbx"))
    assert np.array_equal(w.transform()[2], [2, 2, 2])
    v.SetTransform((0, 0, 0), (2, 2, 2), (0, 0, 0), (0, 0, 0))
    a, b = _render(w), _render(v)
    assert np.array_equal(a, b)
    v.SetTransform((0, 0, 0), (1, 1, 1), (0, 0, 0), (0, 0, 0))
    assert vbx.read_vbx(good)["end"] == len(good)
    # references outside the pools are refused before any kernel walks them:
    # a child slot naming a leaf past the leaf pool, a child list id past the
    # list pool, an unused leaf's atlas corner outside the atlas
    p0 = [v.pool(0, l).copy() for l in range(5)]
    p1 = [v.pool(1, l).copy() if l else None for l in range(5)]
    nleaf = len(p0[0])

    def write(name, q0, q1):
        data = vbx.write_vbx(_meta(v), [x.tobytes() for x in q0], [x.tobytes() if x is not None else b"" for x in q1],
                             [(3, v.atlas())])
        (tmp_path / name).write_bytes(data)
        return str(tmp_path / name)

    c = p1[1].reshape(-1).copy()
    c[np.nonzero(c != np.uint64(2 ** 64 - 1))[0][0]] = np.uint64((nleaf + 5) << 16)
    q1 = list(p1)
    q1[1] = c
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(write("badslot.vbx", p0, q1))
    assert e.value.code == -9
    q0 = list(p0)
    q0[1] = p0[1].copy()
    q0[1]["mChildList"][0] = np.uint64(1 | (1 << 8) | ((len(p1[1]) + 3) << 16))
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(write("badlist.vbx", q0, p1))
    assert e.value.code == -9
    q0 = list(p0)
    q0[0] = p0[0].copy()
    q0[0]["mFlags"][0] = 0
    q0[0]["mValue"][0] = (10 ** 6, 0, 0)
    with pytest.raises(GVDBError) as e:
        w.LoadVBX(write("badcorner.vbx", q0, p1))
    assert e.value.code == -9
    # the unmodified pools still load
    w.LoadVBX(write("same.vbx", p0, p1))
    v.close()
    w.close()
