"""The CPU oracle pinned against brute-force models and closed forms (no GPU).

The reference ships no tests, golden vectors or fixtures for this path, so the
oracle is pinned by what the mathematics fixes: set-based topology models on
small inputs, the closed-form atlas layout, per-point subcell binning, an FP64
brute-force density field, brute-force apron lookups, and closed-form renders
(missed rays, a uniform-density emission-absorption slab, runs of bricks
separated by empty nodes, a row longer than the walk's iteration cap, an
isosurface on an axis through one point, the camera's projection).
"""
import math

import numpy as np
import pytest

import model as M
import oracle as O

TREE = (3, 3, 3, 4, 5)


def _cloud(n, seed, lo=70.0, hi=440.0):
    g = np.random.default_rng(seed)
    return (lo + (hi - lo) * g.random((n, 3))).astype(np.float32)


def _ball(n, seed, centre=(200.0, 180.0, 220.0), radius=30.0):
    g = np.random.default_rng(seed)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius + g.standard_normal(n)
    return (np.asarray(centre) + d * r[:, None]).astype(np.float32)


def _pools_match(o, expected, levels=5):
    for lev in range(levels):
        a = o.pool(0, lev)
        e = expected[lev]
        if e is None:
            assert len(a) == 0, "level %d should be empty" % lev
            continue
        assert len(a) == len(e["mPos"]), "level %d: %d nodes, model %d" % (lev, len(a), len(e["mPos"]))
        for f in ("mLev", "mFlags", "mPos", "mValue", "mParent", "mChildList"):
            assert np.array_equal(a[f], e[f]), "level %d field %s" % (lev, f)
        if lev > 0:
            assert np.array_equal(o.pool(1, lev), e["clists"]), "level %d child lists" % lev


# ------------------------------------------------------------------ topology
@pytest.mark.parametrize("n,seed,radius", [(1, 1, 0.5), (2, 2, 0.5), (40, 3, 0.5), (300, 4, 0.5), (40, 5, 1.0),
                                           (300, 6, 2.0), (120, 7, 3.0), (200, 8, 1.5)])
def test_topology_matches_set_model(n, seed, radius):
    pts = _cloud(n, seed) if n > 2 else (np.array([[100.3, 200.7, 300.1], [350.5, 90.2, 140.9]], np.float32)[:n] + 0)
    m = M.topology_model(pts, radius)
    o = O.Oracle()
    rc = o.rebuild_topology(pts, radius)
    assert rc == 0
    if any(len(m["levels"][l]) == 1 for l in range(m["depth"])):
        # a level below the root with a single node becomes the walk's top
        # node; linking above it is outside the reference's defined behaviour
        assert o.top_lev() < m["depth"]
        return
    assert o.undefined_events() == 0
    assert o.top_lev() == m["depth"]
    _pools_match(o, M.expected_pools(m))


def test_topology_sphere_model_and_invariants():
    pts = _ball(2500, 9)
    m = M.topology_model(pts, 2.0)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    assert o.undefined_events() == 0
    _pools_match(o, M.expected_pools(m))
    # every point lies in a level-0 node
    leaves = {tuple(p) for p in o.pool(0, 0)["mPos"].tolist()}
    for p in np.floor(pts / 8).astype(np.int64) * 8:
        assert tuple(p.tolist()) in leaves


def test_topology_translation_and_negative_points():
    """Points with a negative coordinate after translation get no key."""
    pts = _cloud(200, 10)
    orig = (-120.0, 5.0, 0.0)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 0.5, orig) == 0
    w = pts + np.asarray(orig, np.float32)
    m = M.topology_model(w, 0.5)
    if all(len(m["levels"][l]) > 1 for l in range(m["depth"])):
        _pools_match(o, M.expected_pools(m))
    leaves = {tuple(p) for p in o.pool(0, 0)["mPos"].tolist()}
    want = {tuple(p) for p in (np.floor(w[np.all(w >= 0, 1)] / 8).astype(np.int64) * 8).tolist()}
    assert leaves == want


def test_bounding_box_and_depth():
    pts = _ball(500, 11)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    b = o.bounds()
    assert np.array_equal(b["pos_min"], np.maximum(0, pts.min(0) - np.float32(1)))
    assert np.array_equal(b["pos_max"], np.maximum(0, pts.max(0) + np.float32(2)))
    o.finish_topology()
    leaves = o.pool(0, 0)["mPos"]
    b = o.bounds()
    assert np.array_equal(b["obj_min"], leaves.min(0).astype(np.float32))
    assert np.array_equal(b["obj_max"], (leaves.max(0) + 8).astype(np.float32))


def test_host_activate_space_same_node_set_as_rebuild():
    """RebuildTopologyCPU (host ActivateSpace, insertion order) builds the same
    node set as the GPU rebuild without dilation; parent and child references
    agree with each other."""
    pts = _ball(3000, 12)
    a = O.Oracle()
    assert a.rebuild_topology_cpu(pts) == 0
    b = O.Oracle()
    assert b.rebuild_topology(pts, 0.5) == 0
    for lev in range(4):
        sa = {tuple(p) for p in a.pool(0, lev)["mPos"].tolist()}
        sb = {tuple(p) for p in b.pool(0, lev)["mPos"].tolist()}
        assert sa == sb, "level %d" % lev
    # reference consistency of the host-built pools
    for lev in range(1, 5):
        nodes = a.pool(0, lev)
        if not len(nodes):
            continue
        cl = a.pool(1, lev)
        kids = a.pool(0, lev - 1)
        for i, nd in enumerate(nodes):
            lst = cl[int(nd["mChildList"]) >> 16]
            for slot in np.nonzero(lst != M.ID_UNDEF64)[0]:
                ci = int(lst[slot]) >> 16
                assert int(kids[ci]["mParent"]) == M.elem(0, lev, i)


# --------------------------------------------------------------------- atlas
def test_update_atlas_closed_form():
    pts = _ball(3000, 13)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    o.finish_topology()
    assert o.update_atlas() == 0
    leaves = o.pool(0, 0)
    B = len(leaves)
    assert B > 256          # forces Z growth
    info = o.atlas_info()
    Z = max(1, math.ceil(B / 256))
    assert info["cnt"] == (16, 16, Z) and info["res"] == (160, 160, 10 * Z) and info["used"] == B
    assert np.array_equal(leaves["mValue"], M.atlas_slots(B))
    amap = o.atlas_map()
    assert len(amap) == 256 * Z
    assert np.array_equal(amap["mLeafNode"][:B], np.arange(B))
    assert np.array_equal(amap["mPos"][:B], leaves["mPos"])
    assert np.all(amap["mLeafNode"][B:] == -1) and np.all(amap["mPos"][B:] == -1)


# ------------------------------------------------------------------ subcells
@pytest.mark.parametrize("radius,scs", [(2.0, 4), (0.5, 4), (3.0, 4), (2.0, 2), (1.5, 8)])
def test_subcells_match_per_point_binning(radius, scs):
    pts = _ball(600, 14)
    o = O.Oracle()
    assert o.rebuild_topology(pts, radius) == 0
    o.finish_topology()
    o.update_atlas()
    rc, n = o.insert_points_subcell(pts, radius, scs)
    assert rc == 0
    leaves = [tuple(p) for p in o.pool(0, 0)["mPos"].tolist()]
    m = M.subcell_model(pts, radius, leaves, scs)
    s = o.subcells(n)
    assert n == int(m["cnt"].sum())
    assert np.array_equal(s["cnt"].astype(np.int64), m["cnt"])
    assert np.array_equal(s["off"].astype(np.int64), m["off"])
    assert np.array_equal(s["pos"].astype(np.int64), m["pos"])
    assert np.array_equal(s["nid"].astype(np.int64), m["nid"])
    assert np.array_equal(s["pnt"], m["pnt"])      # canonical (point) order inside each subcell


def test_subcells_skip_nohit_and_negative_points():
    pts = _ball(300, 15)
    extra = np.array([[100.0, 100.0, 1.0e10], [-3.0, 180.0, 200.0]], np.float32)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    o.finish_topology()
    o.update_atlas()
    _, n0 = o.insert_points_subcell(pts, 2.0)
    _, n1 = o.insert_points_subcell(np.concatenate([pts, extra]), 2.0)
    assert n0 == n1


# ------------------------------------------------------------ density, apron
def _gathered(pts, radius, boundval=0.0, precision=64, scs=4):
    o = O.Oracle()
    assert o.rebuild_topology(pts, radius) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts, radius, scs)
    assert rc == 0
    o.gather_density(radius, scs, precision)
    o.update_apron(boundval, precision)
    return o


@pytest.mark.parametrize("radius,scs", [(2.0, 4), (1.5, 4), (3.0, 2), (0.75, 4)])
def test_gather_equals_brute_force_field(radius, scs):
    pts = _ball(1500, 16)
    o = _gathered(pts, radius, scs=scs)
    leaves = o.pool(0, 0)
    lo = leaves["mPos"].min(0)
    hi = leaves["mPos"].max(0) + 8
    g = M.density_grid(pts, radius, lo, hi)
    got = M.brick_interiors(o.atlas(64), leaves["mPos"], leaves["mValue"])
    for i, p in enumerate(leaves["mPos"]):
        want = g[p[0] - lo[0]:p[0] - lo[0] + 8, p[1] - lo[1]:p[1] - lo[1] + 8, p[2] - lo[2]:p[2] - lo[2] + 8]
        assert np.abs(got[i] - want).max() <= 1e-12, "brick %d at %s" % (i, p)


def test_gather_fp32_close_to_fp64():
    pts = _ball(1500, 17)
    a = _gathered(pts, 2.0, precision=32).atlas(32).astype(np.float64)
    b = _gathered(pts, 2.0, precision=64).atlas(64)
    assert np.abs(a - b).max() < 2e-6
    assert np.array_equal(a == 0, b == 0)


# ------------------------------------------------------------- level set
def _levelset(pts, radius, precision=64, scs=4):
    o = O.Oracle()
    assert o.rebuild_topology(pts, radius) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts, radius, scs)
    assert rc == 0
    o.gather_level_set(radius, scs, precision)
    return o


@pytest.mark.parametrize("radius,scs", [(2.0, 4), (1.5, 2)])
def test_level_set_equals_brute_force_over_binned_lists(radius, scs):
    """min(3, min over the points binned into the voxel's subcell of |p-v| - r),
    with the lists from the brute-force binning model."""
    pts = _ball(600, 31)
    o = _levelset(pts, radius, scs=scs)
    leaves = o.pool(0, 0)
    corners = [tuple(int(c) for c in p) for p in leaves["mPos"]]
    sm = M.subcell_model(pts, radius, corners, scs=scs)
    got = M.brick_interiors(o.atlas(64), leaves["mPos"], leaves["mValue"])
    sd = 8 // scs
    for s in range(len(sm["cnt"])):
        b, loc = divmod(s, sd ** 3)
        sx, sy, sz = loc % sd * scs, (loc // sd) % sd * scs, loc // (sd * sd) * scs
        lst = sm["pnt"][sm["off"][s]:sm["off"][s] + sm["cnt"][s]].astype(np.float64)
        for z in range(scs):
            for y in range(scs):
                for x in range(scs):
                    v = np.array([corners[b][0] + sx + x, corners[b][1] + sy + y, corners[b][2] + sz + z], np.float64)
                    want = 3.0
                    if len(lst):
                        want = min(3.0, float(np.min(np.sqrt(((lst - v) ** 2).sum(1)) - np.float64(np.float32(radius)))))
                    assert abs(got[b, sx + x, sy + y, sz + z] - want) <= 1e-12, (s, x, y, z)


def test_level_set_is_minus_density_inside_the_radius():
    """Both gathers pick the nearest binned point: where the density r - c is
    positive the level set c - r is exactly its negative; elsewhere the level
    set is not negative."""
    pts = _ball(1500, 32)
    d = _gathered(pts, 2.0, precision=64)
    ls = _levelset(pts, 2.0, precision=64)
    leaves = d.pool(0, 0)
    a = M.brick_interiors(d.atlas(64), leaves["mPos"], leaves["mValue"])
    b = M.brick_interiors(ls.atlas(64), leaves["mPos"], leaves["mValue"])
    inside = a > 0
    assert inside.any() and np.array_equal(b[inside], -a[inside])
    assert (b[~inside] >= 0).all() and (b <= 3.0).all()


@pytest.mark.parametrize("scs", [4, 2])
def test_level_set_of_an_empty_insert_is_three(scs):
    """An insert that bins no point still has its subcell corners, and the
    minimum over no entries is the starting value 3 on every brick voxel."""
    pts = _ball(600, 34)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts[:0], 2.0, scs)
    assert rc == 0 and n == 0
    sub = o.subcells(0)
    leaves = o.pool(0, 0)
    sd = 8 // scs
    assert len(sub["pos"]) == len(leaves) * sd ** 3
    assert np.array_equal(sub["pos"].reshape(-1, 3)[::sd ** 3], leaves["mPos"].astype(sub["pos"].dtype))
    o.gather_level_set(2.0, scs, 64)
    got = M.brick_interiors(o.atlas(64), leaves["mPos"], leaves["mValue"])
    assert (got == 3.0).all()


def test_level_set_fp32_close_to_fp64():
    pts = _ball(1500, 33)
    a = _levelset(pts, 2.0, precision=32)
    b = _levelset(pts, 2.0, precision=64)
    assert np.abs(a.atlas(32).astype(np.float64) - b.atlas(64)).max() < 4e-6


@pytest.mark.parametrize("boundval", [0.0, -1.5])
def test_apron_matches_neighbour_lookup(boundval):
    pts = _ball(1200, 18)
    o = _gathered(pts, 2.0, boundval=boundval)
    leaves = o.pool(0, 0)
    pos = leaves["mPos"]
    slot = leaves["mValue"] - 1          # corner of the padded 10^3 brick
    lo = pos.min(0) - 8
    hi = pos.max(0) + 16
    g = M.density_grid(pts, 2.0, lo, hi)
    have = {tuple(p) for p in pos.tolist()}
    a = o.atlas(64)
    loc = np.stack(np.meshgrid(np.arange(10), np.arange(10), np.arange(10), indexing="ij"), -1).reshape(-1, 3)
    ap = loc[np.any((loc == 0) | (loc == 9), axis=1)]
    assert len(ap) == 488
    for i in range(len(pos)):
        wv = pos[i] + ap - 1
        got = a[slot[i, 2] + ap[:, 2], slot[i, 1] + ap[:, 1], slot[i, 0] + ap[:, 0]]
        owner = np.floor_divide(wv, 8) * 8
        exists = np.array([tuple(o_) in have for o_ in owner.tolist()])
        want = np.where(exists, g[wv[:, 0] - lo[0], wv[:, 1] - lo[1], wv[:, 2] - lo[2]], boundval)
        assert np.abs(got - want).max() <= 1e-12, "brick %d" % i


# -------------------------------------------------------------------- render
def _scene(w, h, to, angs, dist, **kw):
    s = dict(width=w, height=h, cam=dict(angs=angs, to=to, dist=dist, dolly=1.0, fov=40.0, near=0.1, far=5000.0),
             light=(0.0, 0.0, 0.0), backclr=(0.1, 0.2, 0.3, 1.0), extinct=(-0.1, 1.0, 0.0), steps=(0.5, 16.0, 0.1),
             cutoff=(0.005, 0.01, 0.0), vrange=(0.0, 0.0, 2.0), shadow=(0.0, 1.0, 0.0))
    s.update(kw)
    return s


def test_camera_projection_closed_form():
    """campos = to + dist*(cos b sin a, sin b, cos b cos a); the image corners
    sit at tan(fov/2)/2 horizontally and tan(fov/2)/(2*aspect) vertically off
    the view axis (Camera3D::updateMatricies, src/gvdb_camera.cpp:175-228:
    proj(0,0) = 2n/sx with sx = tan(fov/2) n)."""
    for (w, h), angs, fov in (((64, 48), (30.0, 20.0, 0.0), 40.0), ((100, 100), (-75.0, 50.0, 0.0), 65.0)):
        sc = _scene(w, h, (256.0, 120.0, 256.0), angs, 600.0)
        sc["cam"]["fov"] = fov
        blk = O.scene_block(sc)
        d2r = 3.141592 / 180.0
        a, b = angs[0] * d2r, angs[1] * d2r
        want = np.array([256.0, 120.0, 256.0]) + 600.0 * np.array([math.cos(b) * math.sin(a), math.sin(b),
                                                                    math.cos(b) * math.cos(a)])
        assert np.allclose(blk[2:5], want, atol=2e-3)
        cams, camu, camv = blk[5:8].astype(np.float64), blk[8:11].astype(np.float64), blk[11:14].astype(np.float64)
        fwd = (np.array([256.0, 120.0, 256.0]) - want)
        fwd /= np.linalg.norm(fwd)
        side = np.cross(fwd, [0, 1, 0])
        side /= np.linalg.norm(side)
        up = np.cross(side, fwd)
        t = math.tan(fov * d2r / 2)
        tl = cams
        assert np.isclose(tl @ side / (tl @ fwd), -t / 2, rtol=1e-4)
        assert np.isclose(tl @ up / (tl @ fwd), t / (2 * w / h), rtol=1e-4)
        centre = cams + 0.5 * camu + 0.5 * camv
        assert np.allclose(centre / np.linalg.norm(centre), fwd, atol=1e-5)


def test_linear_transfer_function_closed_form():
    lut = O.linear_tf(0.25, 0.75, (0.0, 0.2, 0.4, 0.0), (1.0, 0.6, 0.4, 1.0))
    n0, n1 = int(0.25 * 16384), int(0.75 * 16384)
    assert np.all(lut[:n0] == 0) and np.all(lut[n1:] == 0)
    x = (np.arange(n0, n1) - n0) / (n1 - n0)
    assert np.allclose(lut[n0:n1, 0], x, atol=1e-6)
    assert np.allclose(lut[n0:n1, 1], 0.2 + 0.4 * x, atol=1e-6)
    assert np.allclose(lut[n0:n1, 2], 0.4, atol=1e-6)


@pytest.mark.parametrize("shading", [7, 4])
def test_render_rays_that_miss_show_background(shading):
    pts = _ball(800, 19)
    o = _gathered(pts, 2.0)
    o.finish_topology()
    sc = _scene(24, 16, (200.0, 180.0, -3000.0), (0.0, 0.0, 0.0), 1000.0)   # looking down -z, volume behind
    img = o.render(O.scene_block(sc), O.linear_tf(0.0, 1.0, (1, 1, 1, 0), (1, 1, 1, 1)), shading, 64)
    back = np.array(sc["backclr"], np.float32)
    if shading == 7:
        back = np.array([0.1, 0.2, 0.3, 0.0], np.float32)     # alpha = 1 - transmittance = 0
    assert np.array_equal(img, np.broadcast_to(back, img.shape))


def _block_points():
    """One point per voxel centre of [56,88) x [64,96) x [64,128): exactly a
    4 x 4 x 8 block of bricks with radius below 1 (no dilation), spread over
    two level-1 nodes."""
    g = np.stack(np.meshgrid(np.arange(56, 88), np.arange(64, 96), np.arange(64, 128), indexing="ij"), -1)
    return (g.reshape(-1, 3) + 0.5).astype(np.float32)


def test_emission_absorption_uniform_slab():
    """Uniform value c in a 64-voxel-deep block: after n samples the
    transmittance is e^n with e = exp(extinct * alpha * step) and the colour
    is col * albedo * (1 - e^n) (rayDeepBrick, kernels/cuda_gvdb_raycast.cuh:
    485-533); the pixel composites over the background with alpha 1 - e^n
    (gvdbRayDeep, kernels/cuda_gvdb_module.cu:60-78).  n is 64 / step up to
    one sample at either end of the lattice.  The ray is slightly oblique: on
    an exactly axis-aligned ray the reference's DDA multiplies an infinite
    tDel by a zero mask (NaN) and stops advancing after the second brick."""
    pts = _block_points()
    o = O.Oracle()
    assert o.rebuild_topology(pts, 0.5) == 0
    o.finish_topology()
    o.update_atlas()
    assert len(o.pool(0, 0)) == 128
    c = 0.8
    o.fill_channel(c)
    step = 0.5
    sc = _scene(1, 1, (72.3, 79.6, 96.0), (0.01, 0.01, 0.0), 300.17, extinct=(-0.05, 0.9, 0.0),
                steps=(step, 16.0, 0.1), vrange=(0.0, 0.0, 2.0))
    col = np.array([1.0, 0.5, 0.25, 1.0], np.float32)
    tf = O.linear_tf(0.0, 1.0, (0, 0, 0, 0), col)
    img = o.render(O.scene_block(sc), tf, 7, 64)[0, 0].astype(np.float64)
    i = int(np.float64(c / 2.0) * 16300)
    tcol = tf[i].astype(np.float64)
    e = math.exp(-0.05 * tcol[3] * step)
    back = np.array(sc["backclr"][:3])
    ok = []
    for n in range(int(64 / step) - 1, int(64 / step) + 2):
        T = e ** n
        clr = np.minimum(tcol[:3] * 0.9 * (1 - T), 1.0)
        want = np.concatenate([back + (1 - T) * (clr - back), [1 - T]])
        ok.append(np.abs(img - want).max())
    assert min(ok) < 1e-6, "pixel %s, closest closed form off by %.3g" % (img, min(ok))   # float RGBA output


def _brick_row(bricks):
    """One point at the centre of each listed brick of the row y, z in [0, 8):
    radius 0.5 activates exactly those bricks."""
    return np.array([[8.0 * k + 4.0, 4.0, 4.0] for k in bricks], np.float32)


def _ea_along_x(o, c, extinct, step):
    """A 1x1 render of the uniform value c down +x along the row y = z ~ 4
    (the ray is 0.01 degree off the axis, see the slab test), entering the
    volume at x = 0.  Returns the pixel and the closed-form pixel of n samples."""
    o.fill_channel(c)
    sc = _scene(1, 1, (95.0, 4.2, 4.2), (-89.99, 0.01, 0.0), 100.17, extinct=(extinct, 0.9, 0.0),
                steps=(step, 16.0, 0.1), vrange=(0.0, 0.0, 2.0))
    col = np.array([1.0, 0.5, 0.25, 1.0], np.float32)
    tf = O.linear_tf(0.0, 1.0, (0, 0, 0, 0), col)
    img = o.render(O.scene_block(sc), tf, 7, 64)[0, 0].astype(np.float64)
    tcol = tf[int(np.float64(c / 2.0) * 16300)].astype(np.float64)
    e = math.exp(extinct * tcol[3] * step)
    back = np.array(sc["backclr"][:3])

    def want(n):
        T = e ** n
        clr = np.minimum(tcol[:3] * 0.9 * (1 - T), 1.0)
        return np.concatenate([back + (1 - T) * (clr - back), [1 - T]])
    return img, want


def _closest_n(img, want, lo, hi):
    errs = {n: np.abs(img - want(n)).max() for n in range(lo, hi + 1)}
    n = min(errs, key=errs.get)
    return n, errs[n]


def test_emission_absorption_skips_empty_nodes():
    """Three 8-brick runs of the row: level-1 nodes 0 and 3 (nodes 1, 2
    empty) and level-1 node 16, in the third level-2 node (the second is
    empty).  The walk steps over the empty nodes at levels 1 and 2 and
    samples each run as the uniform slab: 3 * 64 / step samples, up to one
    at either end of each run (rayCast, kernels/cuda_gvdb_raycast.cuh:
    567-611).  A walk that loses a run after an empty node, or enters a run
    twice, is off by 128 samples or more."""
    runs = [range(0, 8), range(24, 32), range(128, 136)]
    o = O.Oracle()
    assert o.rebuild_topology(_brick_row([k for r in runs for k in r]), 0.5) == 0
    o.finish_topology()
    o.update_atlas()
    assert len(o.pool(0, 0)) == 24 and len(o.pool(0, 1)) == 3
    step = 0.5
    img, want = _ea_along_x(o, 0.8, -0.05, step)
    n0 = int(3 * 64 / step)
    n, err = _closest_n(img, want, n0 - 200, n0 + 200)
    assert abs(n - n0) <= 3 and err < 1e-6, (n, err)


def _bricks_before_cap(top, per_entry):
    """Bricks a walk down the row samples before its 256th iteration when it
    spends top - 1 iterations descending from the root to the first brick,
    per_entry more on every level-1 or level-2 node it enters after that,
    and one per brick."""
    it, K = top - 1, 0
    while True:
        if K and K % 8 == 0:
            it += per_entry             # a new level-1 node
            if K % 64 == 0 and top > 2:
                it += per_entry         # a new level-2 node
        if it >= 256:
            return K
        it += 1
        K += 1


def test_emission_absorption_stops_at_max_iter():
    """A row of 300 bricks (2400 voxels) is longer than the walk's cap of
    256 iterations (kernels/cuda_gvdb_raycast.cuh:24, :567).  An iteration
    either descends into an occupied node or samples one brick; entering a
    node takes one descent, and may take a second where the walk, back at
    the parent level on the boundary, finds the node it just left.  So the
    bricks sampled before the cap lie between the counts with one and with
    two iterations per node entered, and the pixel is the uniform slab of a
    whole number of them (16 samples per brick at step 0.5, up to one at
    either end).  A row of 200 bricks is sampled whole.  No cap, a cap on
    samples or a cap on bricks alone falls outside these bounds."""
    step = 0.5
    for L, capped in ((200, False), (300, True)):
        o = O.Oracle()
        assert o.rebuild_topology(_brick_row(range(L)), 0.5) == 0
        o.finish_topology()
        o.update_atlas()
        assert len(o.pool(0, 0)) == L
        top = o.top_lev()
        hi, lo = _bricks_before_cap(top, 1), _bricks_before_cap(top, 2)
        img, want = _ea_along_x(o, 0.8, -0.0008, step)
        n, err = _closest_n(img, want, 0, int(8 * L / step) + 10)
        K = round(n * step / 8)
        assert abs(n - 8 * K / step) <= 1 and err < 1e-6, (n, err)
        if capped:
            assert hi < L and lo <= K <= hi, (lo, K, hi)
        else:
            assert K == L
        _, cap = o.render(O.scene_block(_scene(1, 1, (95.0, 4.2, 4.2), (-89.99, 0.01, 0.0), 100.17)),
                          O.linear_tf(0.0, 1.0, (0, 0, 0, 0), (1, 1, 1, 1)), 7, 64, capped=True)
        assert bool(cap[0, 0]) == capped


def test_isosurface_on_axis_through_one_point():
    """One point at the integer position P with radius 2: on the line through
    the voxel centres x = P.x + 0.5, y = P.y + 0.5 the trilinear field is
    2 - |z - 0.5 - P.z|, so a ray down -z first reaches 0.5 within one step
    of z = P.z + 2; the normal there is +z and the Phong term is
    0.9 * (l . n) + 0.1 with no shadow (gvdbRaySurfaceTrilinear,
    kernels/cuda_gvdb_module.cu:38-57, :103-120)."""
    P = np.array([200.0, 180.0, 220.0], np.float32)
    pts = np.array([P, [60.0, 60.0, 60.0]], np.float32)     # the second keeps two level-1 nodes
    o = _gathered(pts, 2.0)
    o.finish_topology()
    light = (200.5 + 500.0, 180.5, 222.0 + 500.0)
    sc = _scene(1, 1, (200.5, 180.5, 220.0), (0.0, 0.0, 0.0), 150.1, light=light, steps=(0.25, 16.0, 0.1),
                vrange=(0.5, 0.0, 1.0), shadow=(0.8, 1.0, 0.0))
    img = o.render(O.scene_block(sc), O.linear_tf(0.0, 1.0, (0, 0, 0, 0), (1, 1, 1, 1)), 4, 64)[0, 0]
    vals = []
    for z in np.arange(221.75, 222.0001, 0.0625):
        l = np.array(light) - np.array([200.5, 180.5, z])
        l /= np.linalg.norm(l)
        vals.append(0.9 * l[2] + 0.1)
    assert min(vals) - 1e-4 <= img[0] <= max(vals) + 1e-4
    assert np.allclose(img[:3], img[0], atol=1e-12) and img[3] == 1.0


# ------------------------------------------------ FP16 subcell variants
def _fp16(pts, radius, scs=4, trans=(0.0, 0.0, 0.0), topo_pts=None):
    o = O.Oracle()
    assert o.rebuild_topology(pts if topo_pts is None else topo_pts, radius) == 0
    o.finish_topology()
    o.update_atlas()
    o.clear_channel()
    rc, n = o.insert_points_subcell_fp16(pts, radius, scs, trans)
    assert rc == 0
    return o, n


def _bounds16(o):
    b = o.bounds()
    mn = np.asarray(b["pos_min"], np.float32)
    return mn, np.asarray(b["pos_max"], np.float32) - mn


@pytest.mark.parametrize("scs", [4, 2])
def test_fp16_entries_are_the_quantised_lists(scs):
    """Same lists as InsertPointsSubcell (no translation), each entry
    u16((w - pos_min) / pos_range * 65535); compared as sorted sets."""
    pts = _ball(800, 41)
    o, n = _fp16(pts, 2.0, scs)
    leaves = [tuple(int(c) for c in p) for p in o.pool(0, 0)["mPos"]]
    sm = M.subcell_model(pts, 2.0, leaves, scs=scs)
    assert n == int(sm["cnt"].sum())
    sub = o.subcells(n)
    assert np.array_equal(sub["cnt"], sm["cnt"]) and np.array_equal(sub["off"], sm["off"])
    mn, rg = _bounds16(o)
    got = o.subcells_fp16(n)
    want = np.stack([M.quantize16(sm["pnt"][:, a], mn[a], rg[a]) for a in range(3)], 1)
    for s in range(len(sm["cnt"])):
        a0, a1 = int(sm["off"][s]), int(sm["off"][s] + sm["cnt"][s])
        ga = got[a0:a1][np.lexsort(got[a0:a1].T[::-1])]
        wa = want[a0:a1][np.lexsort(want[a0:a1].T[::-1])]
        assert np.array_equal(ga, wa), s


def test_fp16_level_set_equals_brute_force_over_dequantised_lists():
    """min(3, min over the subcell's entries of |q - v| - r), q the
    correctly rounded float32 u / 65535 * range + min (fused)."""
    pts = _ball(500, 42)
    o, n = _fp16(pts, 2.0)
    o.gather_level_set_fp16(2.0, 4, 64)
    mn, rg = _bounds16(o)
    sub = o.subcells(n)
    q16 = o.subcells_fp16(n)
    deq = np.array([[M.dequantize16(q16[e, a], mn[a], rg[a]) for a in range(3)] for e in range(n)], np.float64)
    leaves = o.pool(0, 0)
    got = M.brick_interiors(o.atlas(64), leaves["mPos"], leaves["mValue"])
    r = float(np.float32(2.0))
    for s in range(len(sub["cnt"])):
        b, loc = divmod(s, 8)
        sx, sy, sz = loc % 2 * 4, (loc // 2) % 2 * 4, loc // 4 * 4
        lst = deq[int(sub["off"][s]):int(sub["off"][s] + sub["cnt"][s])]
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    v = np.array([leaves["mPos"][b][0] + sx + x, leaves["mPos"][b][1] + sy + y,
                                  leaves["mPos"][b][2] + sz + z], np.float64)
                    want = 3.0
                    if len(lst):
                        want = min(3.0, float(np.min(np.sqrt(((lst - v) ** 2).sum(1)) - r)))
                    assert abs(got[b, sx + x, sy + y, sz + z] - want) <= 1e-12, (s, x, y, z)


def test_fp16_level_set_is_within_a_quantisation_step_of_the_float_one():
    pts = _ball(1500, 43)
    o, n = _fp16(pts, 2.0)
    o.gather_level_set_fp16(2.0, 4, 64)
    f = _levelset(pts, 2.0, precision=64)
    mn, rg = _bounds16(o)
    step = float(np.linalg.norm(rg / 65535.0))
    leaves = o.pool(0, 0)
    a = M.brick_interiors(o.atlas(64), leaves["mPos"], leaves["mValue"])
    b = M.brick_interiors(f.atlas(64), leaves["mPos"], leaves["mValue"])
    d = np.abs(a - b)
    assert d.max() <= step and d.max() > 0


def test_fp16_binning_rule_differs_from_the_float_one_only_under_translation():
    """InsertPointsSubcell_FP16 tests p + trans (for < 0 and NOHIT) in both
    passes; the float variant's insert pass also drops p < 0.  Points with
    p < 0 <= p + trans are binned by the first only."""
    base = _ball(600, 44)
    t = np.array([5.0, 3.0, 0.0], np.float32)
    q = (base - t).astype(np.float32)
    q[:50, 0] = np.float32(-0.5)                  # p.x < 0 <= w.x = 4.5
    q[50:60, 2] = np.float32(1.0e10)              # NOHIT: skipped by both
    w = (q + t).astype(np.float32)
    o16, n16 = _fp16(q, 2.0, trans=t, topo_pts=np.concatenate([base, w[:50]]))
    leaves = [tuple(int(c) for c in p) for p in o16.pool(0, 0)["mPos"]]
    keep16 = (q[:, 2] != np.float32(1.0e10)) & np.all(w >= 0, 1)
    keep32 = keep16 & np.all(q >= 0, 1)
    sm16 = M.subcell_model(w[keep16], 2.0, leaves, scs=4)
    sm32 = M.subcell_model(w[keep32], 2.0, leaves, scs=4)
    assert np.array_equal(o16.subcells(n16)["cnt"], sm16["cnt"])
    assert o16.undefined_events() == 0
    o32 = O.Oracle()
    assert o32.rebuild_topology(np.concatenate([base, w[:50]]), 2.0) == 0
    o32.finish_topology()
    o32.update_atlas()
    o32.clear_channel()
    rc, n32 = o32.insert_points_subcell(q, 2.0, 4, t)
    assert rc == 0 and np.array_equal(o32.subcells(n32)["cnt"], sm32["cnt"])
    assert n16 > n32


# ------------------------------------------------------ incremental topology
def _tree_consistent(o, nlev=5):
    """Every filled child slot names a node whose parent names the slot's
    owner, and the child's corner lies in the slot's cell."""
    res, rng = M.tree_ranges((3, 3, 3, 4, 5))
    for lev in range(1, nlev):
        nodes = o.pool(0, lev)
        if not len(nodes):
            continue
        cl = o.pool(1, lev)
        kids = o.pool(0, lev - 1)
        for i, nd in enumerate(nodes):
            lst = cl[int(nd["mChildList"]) >> 16]
            for b in np.nonzero(lst != M.ID_UNDEF64)[0]:
                ci = int(lst[b]) >> 16
                assert int(kids[ci]["mParent"]) == M.elem(0, lev, i)
                c = (kids[ci]["mPos"] - nd["mPos"]) * res[lev] // rng[lev]
                assert (int(c[2]) * res[lev] + int(c[1])) * res[lev] + int(c[0]) == b


@pytest.mark.parametrize("shift,accum", [((15.0, 0.0, 0.0), False), ((70.0, 0.0, 0.0), False),
                                         ((15.0, 0.0, 0.0), True), ((70.0, -40.0, 0.0), True)])
def test_incremental_topology_matches_model(shift, accum):
    """RebuildTopology on points A, then the incremental build on points B
    (the third RebuildTopology call after Configure, or AccumulateTopology):
    pools, child lists and the leaf used count equal the dictionary model."""
    A = _ball(400, 51)
    B = _ball(400, 52, centre=(200.0 + shift[0], 180.0 + shift[1], 220.0 + shift[2]))
    o = O.Oracle()
    assert o.rebuild_topology(A, 2.0) == 0
    if not accum:
        assert o.rebuild_topology(A, 2.0) == 0      # depth change from -1: full again, sets mCurrDepth
    P, C = M.pools_as_dicts(o)
    rc = o.accumulate_topology(B, 2.0) if accum else o.rebuild_topology(B, 2.0)
    assert rc == 0 and o.undefined_events() == 0
    P2, C2, used = M.incremental_model(P, C, B, 2.0, 2, accum)
    Q, D = M.pools_as_dicts(o)
    assert used is not None and len(P2[0]) > len(P[0])
    assert Q == P2 and D == C2
    assert o.pool_used(0, 0) == used
    _tree_consistent(o)
    leaves = o.pool(0, 0)
    flagged = {tuple(p) for p in leaves["mPos"][leaves["mFlags"] == 1].tolist()}
    assert flagged
    if not accum:
        # every flagged leaf is a brick of some point's box [int(w) - r, int(w) + r]
        want = set()
        for p in B:
            lo = [int(np.float32(int(p[a]) - np.float32(2.0))) // 8 * 8 for a in range(3)]
            hi = [int(np.float32(int(p[a]) + np.float32(2.0))) // 8 * 8 for a in range(3)]
            for x in range(lo[0], hi[0] + 1, 8):
                for y in range(lo[1], hi[1] + 1, 8):
                    for z in range(lo[2], hi[2] + 1, 8):
                        want.add((x, y, z))
        assert flagged == want


def _full(pts, r):
    o = O.Oracle()
    assert o.rebuild_topology(pts, r) == 0
    return o


def test_incremental_topology_without_new_nodes_changes_nothing():
    """Points inside the existing leaves only: the reference returns before
    applying the markers, so even unused leaves stay flagged."""
    A = _ball(400, 53)
    o = O.Oracle()
    assert o.rebuild_topology(A, 2.0) == 0
    assert o.rebuild_topology(A, 2.0) == 0
    before = [o.pool(0, l).copy() for l in range(3)]
    assert o.rebuild_topology(A[:50], 2.0) == 0
    after = [o.pool(0, l) for l in range(3)]
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_second_rebuild_after_configure_is_full():
    """mCurrDepth starts at -1 and only the depth-change branch sets it: the
    second RebuildTopology is a full rebuild, the third incremental."""
    A = _ball(400, 54)
    B = _ball(400, 55, centre=(215.0, 180.0, 220.0))
    o = O.Oracle()
    assert o.rebuild_topology(A, 2.0) == 0
    assert o.rebuild_topology(B, 2.0) == 0
    f = _full(B, 2.0)
    assert all(np.array_equal(o.pool(0, l), f.pool(0, l)) for l in range(3))
    o.request_full_rebuild()
    assert o.rebuild_topology(A, 2.0) == 0
    assert all(np.array_equal(o.pool(0, l), _full(A, 2.0).pool(0, l)) for l in range(3))


# ------------------------------------------------------------------ VBX file
def test_vbx_field_sizes_follow_the_specification():
    """Byte counts of GVDB_FILESPEC.txt: header 1+1+4*12+4, grid header
    256+1+1+1+12+4+12+4+4+8+1+4+1+12+12 (SaveVBX's order: channel count
    before the atlas size), topology 4+8, per level 4+4+12+4+4+4+4."""
    import vbx
    assert vbx.HEADER.size == 54 and vbx.GRID.size == 333 and vbx.TOPO.size == 12 and vbx.LEVEL.size == 36


def _vbx_meta(o, bricks):
    res, rng = M.tree_ranges((3, 3, 3, 4, 5))
    return dict(levels=5, root=M.elem(0, 2, 0), log2dims=(3, 3, 3, 4, 5), res=res, ranges=rng, apron=1, brick=8,
                axiscnt=o.atlas_info()["cnt"], bricks=bricks)


def test_vbx_round_trip_of_oracle_volume():
    import vbx
    pts = _ball(800, 71)
    o = _gathered(pts, 2.0)
    p0 = [o.pool(0, l).tobytes() for l in range(5)]
    p1 = [o.pool(1, l).tobytes() if l else b"" for l in range(5)]
    a = o.atlas(32)
    buf = vbx.write_vbx(_vbx_meta(o, len(o.pool(0, 0))), p0, p1, [(3, a)])
    r = vbx.read_vbx(buf)
    assert r["major"] == 1 and r["minor"] == 11 and r["offsets"] == (62,) and r["end"] == len(buf)
    assert r["dtype"] == "f" and r["topology"] == 2 and r["layout"] == 0 and r["channels"] == 1
    assert r["pools0"] == p0 and r["pools1"][1:] == p1[1:] and np.array_equal(r["atlases"][0][1], a)
    assert r["atlas_bytes"] == a.nbytes and r["axisres"] == tuple(int(x) for x in o.atlas_info()["res"])
    assert len(buf) == 62 + 333 + 12 + 5 * 36 + sum(map(len, p0)) + sum(map(len, p1)) + 8 + a.nbytes


# ---------------------------------------------------------- Compute operators
def test_hash_is_the_one_at_a_time_hash():
    """Applied to a one-byte key the reference's hash is Jenkins'
    one-at-a-time hash, whose published value for "a" is 0xca2e9442."""
    assert O.lib().orc_hash(0x61) == 0xCA2E9442
    vals = np.array([O.lib().orc_random3(float(x), float(y), 3.0) for x in range(40) for y in range(40)])
    assert vals.min() >= 0.0 and vals.max() < 1.0 and abs(vals.mean() - 0.5) < 0.02


def _computed(field, effect, iters, params, precision=64, skip=True, update_apron=False):
    pts = _ball(600, 91)
    o = _gathered(pts, 2.0, precision=precision)
    info = o.atlas_info()
    rx, ry, rz = info["res"]
    z, y, x = np.meshgrid(np.arange(rz), np.arange(ry), np.arange(rx), indexing="ij")
    a = field(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    o.set_atlas(a, precision)
    assert o.compute(effect, iters, params, update_apron, skip, 0.0, precision) == 0
    return o, a, o.atlas(precision)


def _interiors(o, a):
    import model as Mm
    info = o.atlas_info()
    used = info["used"]
    cx, cy = info["cnt"][0], info["cnt"][1]
    m = np.zeros(a.shape, bool)
    for s in range(used):
        sx, sy, sz = s % cx, (s // cx) % cy, s // (cx * cy)
        m[sz * 10 + 1:sz * 10 + 9, sy * 10 + 1:sy * 10 + 9, sx * 10 + 1:sx * 10 + 9] = True
    return m


def test_smooth_keeps_linear_fields_and_adds_p2():
    """The six-neighbour average of a linear field is the field, so one pass
    of (p1 v + neighbours) / (p1 + 6) + p2 adds exactly p2 on every brick
    interior and leaves the other voxels alone."""
    f = lambda x, y, z: 0.5 * x - 0.25 * y + 0.125 * z + 3.0
    o, a, b = _computed(f, 113, 1, (2.0, 0.75, 0.0))
    m = _interiors(o, a)
    assert np.abs(b[m] - (a[m] + 0.75)).max() < 1e-9
    assert np.array_equal(b[~m], a[~m])


def test_smooth_of_a_spike_spreads_by_the_weights():
    """A unit spike at one interior voxel: the voxel keeps p1 / (p1 + 6), each
    face neighbour 1 / (p1 + 6)."""
    def f(x, y, z):
        return ((x == 15) & (y == 15) & (z == 15)).astype(np.float64)
    o, a, b = _computed(f, 113, 1, (3.0, 0.0, 0.0))
    assert abs(b[15, 15, 15] - 3.0 / 9.0) < 1e-15
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        assert abs(b[15 + d[2], 15 + d[1], 15 + d[0]] - 1.0 / 9.0) < 1e-15


def test_grow_and_fill_closed_forms():
    f = lambda x, y, z: np.where(x % 3 == 0, 0.0, np.where(x % 3 == 1, 0.005, 1.0))
    o, a, b = _computed(f, 115, 1, (0.5, 0.0, 0.0))
    m = _interiors(o, a)
    want = np.where(a == 0, 0.0, a + 5.0)
    want = np.where(want < 0.01, 0.0, want)
    assert np.array_equal(b[m], want[m]) and np.array_equal(b[~m], a[~m])
    o, a, b = _computed(lambda x, y, z: x * 0.0, 110, 1, (2.5, 0.0, 0.0), skip=False)
    rx = b.shape[2]
    assert (b[1:, 1:, 1:] == 2.5).all() and (b[0] == 0).all() and (b[:, 0] == 0).all() and (b[:, :, 0] == 0).all()
    o, a, b = _computed(lambda x, y, z: x * 0.0, 110, 1, (2.5, 0.0, 0.0), skip=True)
    hi = [min(8 * ((c * 8 + 7) // 8), r - 1) for c, r in zip(o.atlas_info()["cnt"], o.atlas_info()["res"])]
    assert (b[1:hi[2] + 1, 1:hi[1] + 1, 1:hi[0] + 1] == 2.5).all()
    assert (b != 0).sum() == hi[0] * hi[1] * hi[2]


# ------------------------------------------------------------ colour channel
def _colored(pts, radius, clr, scs=4):
    o = O.Oracle()
    assert o.rebuild_topology(pts, radius) == 0
    o.finish_topology()
    o.update_atlas()
    o.add_color_channel()
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts, radius, scs, colors=clr)
    assert rc == 0
    return o, n


def _subcell_voxels(o, scs=4):
    """(subcell index, atlas indices (z, y, x) of its voxels) for every subcell."""
    leaves = o.pool(0, 0)
    sd = 8 // scs
    for b, lf in enumerate(leaves):
        for s in range(sd ** 3):
            lx, ly, lz = s % sd * scs, s // sd % sd * scs, s // sd // sd * scs
            v = lf["mValue"]
            zz, yy, xx = np.meshgrid(np.arange(scs) + v[2] + lz, np.arange(scs) + v[1] + ly, np.arange(scs) + v[0] + lx,
                                     indexing="ij")
            yield b * sd ** 3 + s, (zz.ravel(), yy.ravel(), xx.ravel())


@pytest.mark.parametrize("accumulate", [False, True])
def test_color_gather_is_the_exact_mean_of_each_list(accumulate):
    """gvdbGatherDensity's colour (particles.cuh:717-731) taken exactly: every
    voxel of a subcell gets (255 [alpha] + sum of its list's bytes) / count,
    or (old + sum) / (count + 1) when accumulating; an empty list stores
    (0, 0, 0, 255).  Lists from the per-point binning model."""
    pts = _ball(700, 61)
    g = np.random.default_rng(7)
    clr = g.integers(0, 1 << 32, len(pts), dtype=np.uint64).astype(np.uint32)
    o, n = _colored(pts, 2.0, clr)
    leaves = [tuple(p) for p in o.pool(0, 0)["mPos"].tolist()]
    m = M.subcell_model(pts, 2.0, leaves, 4)
    old = None
    if accumulate:
        old = g.integers(0, 1 << 32, o.color_atlas().shape, dtype=np.uint64).astype(np.uint32)
        o.set_color_atlas(old)
    assert o.gather_color(4, accumulate, exact=True) == 0
    a = o.color_atlas()
    byte = (clr[:, None] >> np.arange(0, 32, 8, dtype=np.uint32)) & 0xFF
    for s, vox in _subcell_voxels(o):
        ids = m["idx"][s]
        tot = byte[ids].astype(np.int64).sum(0) if ids else np.zeros(4, np.int64)
        got = (a[vox][:, None] >> np.arange(0, 32, 8, dtype=np.uint32)) & 0xFF
        if accumulate:
            ob = (old[vox][:, None] >> np.arange(0, 32, 8, dtype=np.uint32)) & 0xFF
            want = (ob.astype(np.int64) + tot[None, :]) // (len(ids) + 1)
        elif not ids:
            want = np.tile(np.array([0, 0, 0, 255]), (len(vox[0]), 1))
        else:
            want = np.tile(np.minimum((tot + np.array([0, 0, 0, 255])) // len(ids), 255), (len(vox[0]), 1))
        assert np.array_equal(got.astype(np.int64), want), s


def test_color_gather_closed_forms_and_the_float_formula():
    """One colour for every point: each byte of a non-empty subcell is that
    byte (alpha: min(255, (255 + n a) / n)); the reference's float sum
    (list order) differs from the exact mean by at most one unit."""
    pts = _ball(900, 62)
    c = np.uint32(0x50C81E0A)      # r 10, g 30, b 200, a 80
    o, n = _colored(pts, 2.0, np.full(len(pts), c, np.uint32))
    assert o.gather_color(4, False, exact=True) == 0
    a = o.color_atlas()
    s = o.subcells(n)
    for sc, vox in _subcell_voxels(o):
        k = int(s["cnt"][sc])
        want = 0xFF000000 if k == 0 else (0x00C81E0A | (min(255, (255 + k * 80) // k) << 24))
        assert np.all(a[vox] == want), sc
    g = np.random.default_rng(8)
    clr = g.integers(0, 1 << 32, len(pts), dtype=np.uint64).astype(np.uint32)
    for acc in (False, True):
        o, n = _colored(pts, 2.0, clr)
        o.gather_color(4, acc, exact=True)
        e = o.color_atlas().view(np.uint8).astype(np.int64)
        o.clear_color()
        o.gather_color(4, acc, exact=False)
        f = o.color_atlas().view(np.uint8).astype(np.int64)
        assert np.abs(e - f).max() <= 1 and np.count_nonzero(e != f) < 0.05 * np.count_nonzero(e)


def test_color_apron_moves_the_same_voxels_as_the_float_apron():
    """UpdateApron<uchar4> walks as UpdateApron<float>: with the float
    channel's bits copied into the colour channel, the two aprons give the
    same bits; the boundary value becomes a truncated, clamped byte."""
    pts = _ball(1000, 63)
    o, n = _colored(pts, 2.0, np.zeros(len(pts), np.uint32))
    o.gather_density(2.0, 4, 32)
    o.set_color_atlas(o.atlas(32).view(np.uint32))
    o.update_apron(0.0, 32)
    o.update_apron_color(0.0)
    assert np.array_equal(o.color_atlas(), o.atlas(32).view(np.uint32))
    for bv, byte in ((3.7, 3), (-2.0, 0), (300.0, 255), (255.0, 255)):
        o.update_apron(bv, 32)
        o.update_apron_color(bv)
        f, c = o.atlas(32), o.color_atlas()
        m = f == np.float32(bv)
        assert m.any()
        assert np.all(c[m] == np.uint32(byte * 0x01010101))


def test_render_color_channel_scales_by_a_uniform_color():
    """SetColorChannel with every voxel one colour h: the emission-absorption
    colour and the isosurface colour are h times the ones without it, per
    channel (the reference multiplies each sample's term, and the surface
    colour, by getColorF)."""
    pts = _ball(1500, 64)
    o = _gathered(pts, 2.0, precision=32)
    o.add_color_channel()
    rx, ry, rz = o.atlas_info()["res"]
    h = np.array([0.2, 0.6, 1.0])
    o.set_color_atlas(np.full((rz, ry, rx), np.uint32(0xFF000000 | (255 << 16) | (153 << 8) | 51), np.uint32))
    lo, hi = pts.min(0), pts.max(0)
    for shading, tf, vr in ((7, O.linear_tf(0.0, 1.0, (1, 1, 1, 0), (1, 0.9, 0.7, 1)), (0.0, 0.0, 2.0)),
                            (4, O.linear_tf(0.0, 1.0, (0, 0, 0, 0), (1, 1, 1, 0.1)), (0.5, 0.0, 1.0))):
        sc = _scene(96, 80, tuple((lo + hi) / 2), (30.0, 20.0, 0.0), 300.0, vrange=vr, light=(400.0, 500.0, 300.0))
        blk = O.scene_block(sc)
        o.set_color_channel(False)
        a = o.render(blk, tf, shading, 32).astype(np.float64)
        o.set_color_channel(True)
        b = o.render(blk, tf, shading, 32).astype(np.float64)
        back = np.asarray(sc["backclr"], np.float64)
        if shading == 7:
            # res = back + alpha (clr - back): clr scales by h
            al = a[..., 3:4]
            clr_a = np.where(al > 0, back[:3] + (a[..., :3] - back[:3]) / np.where(al > 0, al, 1), 0)
            want = back[:3] + al * (h * clr_a - back[:3])
            hitm = al[..., 0] > 1e-3
        else:
            want = a[..., :3] * h
            hitm = np.any(np.abs(a[..., :3] - back[:3]) > 1e-6, axis=-1)
        assert hitm.sum() > 100
        assert np.abs(b[..., :3][hitm] - want[hitm]).max() < 2e-5, shading
        assert np.array_equal(b[..., 3], a[..., 3])
