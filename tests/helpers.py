"""Shared drivers for the parity tests: the same seeded inputs through the HIP
library (via the C ABI) and through the CPU oracle, plus comparison helpers."""
import numpy as np

TREE = (5, 4, 3, 3, 3)          # Configure(q4..q0): leaf-first log2 dims {3,3,3,4,5}
ZERO3 = (0.0, 0.0, 0.0)
NODE_FIELDS = ("mLev", "mFlags", "mPos", "mValue", "mParent", "mChildList")   # fields the path writes


def gpu_volume(g, pts, radius=2.0, host_topology=False, scs=4, atlas_type=None, boundval=0.0, axiscnt=(16, 16, 1),
               stop_after=None):
    v = g.VolumeGVDB(0)
    v.Configure(*TREE)
    v.SetChannelDefault(*axiscnt)
    v.AddChannel(0, g.T_FLOAT if atlas_type is None else atlas_type, 1)
    v.SetPoints(pts)
    if host_topology:
        v.RebuildTopologyCPU(pts)
        v.FinishTopology(True, True)
    else:
        v.RequestFullRebuild(True)
        v.RebuildTopology(len(pts), radius, ZERO3)
        v.FinishTopology(False, True)
    if stop_after == "topology":
        return v, 0
    v.UpdateAtlas()
    v.ClearChannel(0)
    n = v.InsertPointsSubcell(scs, len(pts), radius, ZERO3)
    if stop_after == "insert":
        return v, n
    v.GatherDensity(scs, len(pts), radius, ZERO3, n, 0, 0, False)
    v.UpdateApron(0, boundval)
    v.synchronize()
    return v, n


def oracle_volume(O, pts, radius=2.0, host_topology=False, scs=4, boundval=0.0, precision=64, axiscnt=(16, 16, 1),
                  stop_after=None):
    o = O.Oracle(axiscnt=axiscnt)
    if host_topology:
        rc = o.rebuild_topology_cpu(pts)
    else:
        rc = o.rebuild_topology(pts, radius)
    assert rc == 0
    o.finish_topology()
    if stop_after == "topology":
        return o, 0
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts, radius, scs)
    assert rc == 0
    if stop_after == "insert":
        return o, n
    for prec in ((32, 64) if precision == "both" else (precision,)):
        o.gather_density(radius, scs, prec)
        o.update_apron(boundval, prec)
    return o, n


def assert_pools_equal(v, o, levels=5):
    for lev in range(levels):
        a, b = v.pool(0, lev), o.pool(0, lev)
        assert len(a) == len(b), "level %d node count %d vs %d" % (lev, len(a), len(b))
        for f in NODE_FIELDS:
            assert np.array_equal(a[f], b[f]), "level %d field %s differs" % (lev, f)
        if lev > 0:
            ca, cb = v.pool(1, lev), o.pool(1, lev)
            assert ca.shape == cb.shape, "level %d child lists %s vs %s" % (lev, ca.shape, cb.shape)
            assert np.array_equal(ca, cb), "level %d child lists differ" % lev


def canonical_nodes(vol, levels=5):
    """(lev, x, y, z) tuples of every node, sorted: topology after canonical ordering."""
    out = []
    for lev in range(levels):
        p = vol.pool(0, lev)
        out += [(lev, int(x), int(y), int(z)) for x, y, z in p["mPos"]]
    return sorted(out)


def subcell_lists(cnt, off, pnt):
    """Per-subcell point lists sorted lexicographically (entries are unordered within a subcell)."""
    res = []
    for c, s in zip(cnt.astype(np.int64), off.astype(np.int64)):
        blk = pnt[s:s + c]
        if c:
            blk = blk[np.lexsort((blk[:, 2], blk[:, 1], blk[:, 0]))]
        res.append(blk)
    return res


def used_brick_mask(atlas_shape, amap, cnt, br=10):
    """Boolean mask of atlas voxels belonging to used slots."""
    rz, ry, rx = atlas_shape
    m = np.zeros(atlas_shape, bool)
    used = np.nonzero(amap["mLeafNode"] >= 0)[0]
    for s in used:
        sx, sy, sz = s % cnt[0], (s // cnt[0]) % cnt[1], s // (cnt[0] * cnt[1])
        m[sz * br:(sz + 1) * br, sy * br:(sy + 1) * br, sx * br:(sx + 1) * br] = True
    return m


def interior_mask(atlas_shape, amap, cnt, br=10, apron=1):
    """Boolean mask of the interior (apron excluded) voxels of used slots."""
    m = np.zeros(atlas_shape, bool)
    used = np.nonzero(amap["mLeafNode"] >= 0)[0]
    for s in used:
        sx, sy, sz = s % cnt[0], (s // cnt[0]) % cnt[1], s // (cnt[0] * cnt[1])
        m[sz * br + apron:(sz + 1) * br - apron, sy * br + apron:(sy + 1) * br - apron,
          sx * br + apron:(sx + 1) * br - apron] = True
    return m


def density_errors(a, b):
    """max |a-b|, max relative error over |b| > 0, and mismatched exact zeros."""
    a = a.astype(np.float64)
    d = np.abs(a - b)
    nz = b != 0
    rel = (d[nz] / np.abs(b[nz])).max() if nz.any() else 0.0
    return d.max() if d.size else 0.0, rel, int(np.count_nonzero((b == 0) & (a != 0)))


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def used_bricks(atlas, amap, cnt, br=10):
    """Yields (slot ids, padded bricks (k, br, br, br)) of the used slots, one
    layer of atlas slots (fixed slot z) at a time, so full-size atlases are
    compared without a whole-atlas mask."""
    c0, c1 = int(cnt[0]), int(cnt[1])
    per = c0 * c1
    leaf = amap["mLeafNode"]
    for sz in range(int(cnt[2])):
        used = np.nonzero(leaf[sz * per:(sz + 1) * per] >= 0)[0]
        if used.size == 0:
            continue
        layer = atlas[sz * br:(sz + 1) * br].reshape(br, c1, br, c0, br).transpose(1, 3, 0, 2, 4)
        yield sz * per + used, layer.reshape(per, br, br, br)[used]


def brick_density_errors(a, b, amap, cnt):
    """density_errors over the used bricks of two full atlases (3D layout)."""
    dmax, rel, zeros = 0.0, 0.0, 0
    for (_, x), (_, y) in zip(used_bricks(a, amap, cnt), used_bricks(b, amap, cnt)):
        d, r, z = density_errors(x, y.astype(np.float64))
        dmax, rel, zeros = max(dmax, d), max(rel, r), zeros + z
    return dmax, rel, zeros


def half_errors(a, b, amap, cnt):
    """FP16 storage against the FP64 field over the used bricks: the largest
    error in units of the half rounding bound (2^-11 |b| + 2^-25: relative
    half-ulp of a normal half, half the subnormal spacing below), and the
    voxels where b is exactly zero and a is not."""
    worst, zeros = 0.0, 0
    for (_, x), (_, y) in zip(used_bricks(a, amap, cnt), used_bricks(b, amap, cnt)):
        x = x.astype(np.float64)
        bound = np.abs(y) * 2.0 ** -11 + 2.0 ** -25
        worst = max(worst, float((np.abs(x - y) / bound).max()))
        zeros += int(np.count_nonzero((y == 0) & (x != 0)))
    return worst, zeros
