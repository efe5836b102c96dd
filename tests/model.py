"""Brute-force models of the point-to-volume path, used to pin the oracle.

These are written from the reference's rules as sets and dense grids, not as
the oracle's sort/scan/walk code, so a slip in either shows up as a mismatch:

  topology  the node set of each level is {covering node of every point}
            (gvdbCalcBrickId, kernels/cuda_gvdb_particles.cuh:352-382) plus,
            for radius >= 1, the nodes of the per-axis box
            [int(p)-r, int(p)+2r] (gvdbCalcExtraBrickId, :467-512); pools hold
            the covering nodes in ascending (x, y, z) order followed by the
            dilation-only nodes in the same order (FindUniqueBrick then
            ActivateExtraBricksGPU, src/gvdb_volume_gvdb.cpp:843-1016); the root
            comes from DetermineDepth (:821-837) on the padded bounding box
            (GetBoundingBox, :5404-5467).
  atlas     slot i = i-th used leaf in pool order, placed x fastest in a
            (16, 16, Z) grid of 10^3 bricks (UpdateAtlas :2630-2704,
            getAtlasPos src/gvdb_allocator.cpp:705-715).
  subcells  a point goes to every subcell whose corner s satisfies
            int(p-r)/sc*sc <= s <= int(p+1+r)/sc*sc on each axis and lies in
            a brick (CountSubcell/InsertSubcell, particles.cuh:218-315).
  density   for points with p - r >= 0 on every axis, every voxel within r of
            a point lies in one of that point's subcells, so the gathered
            value (gvdbGatherDensity, particles.cuh:696-733) is the plain
            max(0, max_p (r - |p - v|)) over all points, at every voxel of an
            existing brick.
  apron     apron voxel at world g = brick corner + local - 1 holds the
            density at g when a brick covers g, else the boundary value
            (gvdbUpdateApronF, kernels/cuda_gvdb_operators.cuh:72-131).
"""
import numpy as np

ID_UNDEFL = 0xFFFFFFFF
ID_UNDEF64 = 0xFFFFFFFFFFFFFFFF


def elem(grp, lev, idx):
    return grp | (lev << 8) | (idx << 16)


def ctrunc_div(a, b):
    """C integer division (truncation toward zero) of int arrays/ints by b > 0."""
    a = np.asarray(a, np.int64)
    q = np.abs(a) // b
    return np.where(a < 0, -q, q)


def tree_ranges(log2dims):
    res = [1 << d for d in log2dims]
    rng = []
    acc = 1
    for r in res:
        acc *= r
        rng.append(acc)
    return res, rng


def topology_model(pts, radius, log2dims=(3, 3, 3, 4, 5)):
    """Expected pools of RebuildTopology on points `pts` (no translation).
    Returns dict(depth, root, levels=[list of (x,y,z) corners per level])."""
    w = np.asarray(pts, np.float32)
    res, rng = tree_ranges(log2dims)
    pmin = np.maximum(np.float32(0), w.min(0) - np.float32(1))
    pmax = np.maximum(np.float32(0), w.max(0) + np.float32(2))
    imin, imax = pmin.astype(np.int64), pmax.astype(np.int64)
    depth = None
    for lev in range(len(res)):
        if np.array_equal(imin // rng[lev], imax // rng[lev]):
            depth = lev
            break
    assert depth is not None and depth > 0
    root = tuple(int(x) for x in (imin // rng[depth]) * rng[depth])
    valid = w[np.all(w >= 0, axis=1)]
    covering = []
    for lev in range(depth):
        k = np.floor(valid / np.float32(rng[lev])).astype(np.int64)
        covering.append(set(map(tuple, k.tolist())))
    extra = [set() for _ in range(depth)]
    if radius >= 1:
        iw = valid.astype(np.int64).astype(np.float32)
        bmin = (iw - np.float32(radius)).astype(np.int64)
        bmax = (iw + np.float32(2) * np.float32(radius)).astype(np.int64)
        for lev in range(depth):
            lo = ctrunc_div(bmin, rng[lev])
            hi = ctrunc_div(bmax, rng[lev])
            for a, b in zip(lo.tolist(), hi.tolist()):
                for x in range(a[0], b[0] + 1):
                    for y in range(a[1], b[1] + 1):
                        for z in range(a[2], b[2] + 1):
                            if (x, y, z) not in covering[lev]:
                                extra[lev].add((x, y, z))
    levels = []
    for lev in range(depth):
        keys = sorted(covering[lev]) + sorted(extra[lev])
        levels.append([(x * rng[lev], y * rng[lev], z * rng[lev]) for x, y, z in keys])
    levels.append([root])
    return dict(depth=depth, root=root, levels=levels, res=res, rng=rng, ncover=[len(c) for c in covering])


def expected_pools(model, nlevels=5):
    """Node fields and child lists implied by a topology model."""
    res, rng, levels, depth = model["res"], model["rng"], model["levels"], model["depth"]
    index = [{p: i for i, p in enumerate(lv)} for lv in levels]
    out = []
    for lev in range(nlevels):
        if lev > depth:
            out.append(None)
            continue
        pos = levels[lev]
        n = len(pos)
        parent = np.full(n, ID_UNDEFL, np.uint64)
        if lev < depth:
            for i, p in enumerate(pos):
                pp = tuple((c // rng[lev + 1]) * rng[lev + 1] for c in p)
                parent[i] = elem(0, lev + 1, index[lev + 1][pp])
        child = np.array([elem(1, lev, i) if lev > 0 else ID_UNDEFL for i in range(n)], np.uint64)
        clists = None
        if lev > 0:
            r = res[lev]
            clists = np.full((n, r ** 3), ID_UNDEF64, np.uint64)
            for j, c in enumerate(levels[lev - 1]):
                pp = tuple((x // rng[lev]) * rng[lev] for x in c)
                i = index[lev][pp]
                cx, cy, cz = ((c[a] - pp[a]) // rng[lev - 1] for a in range(3))
                clists[i, (cz * r + cy) * r + cx] = elem(0, lev - 1, j)
        out.append(dict(mLev=np.full(n, lev, np.uint8), mFlags=np.ones(n, np.uint8),
                        mPos=np.asarray(pos, np.int32).reshape(n, 3), mValue=np.full((n, 3), -1, np.int32),
                        mParent=parent, mChildList=child, clists=clists))
    return out


def atlas_slots(nbricks, cnt_xy=(16, 16), br=10, apron=1):
    i = np.arange(nbricks)
    per = cnt_xy[0] * cnt_xy[1]
    return np.stack([(i % cnt_xy[0]) * br + apron, (i // cnt_xy[0]) % cnt_xy[1] * br + apron, (i // per) * br + apron], 1)


def subcell_model(pts, radius, leaves, scs=4, res0=8):
    """Expected subcell tables: leaves = list of brick corners in pool order,
    all linked.  Returns cnt, off, pos, nid, the entry list in point order and
    `idx`, each subcell's list as point indices."""
    w = np.asarray(pts, np.float32)
    sd = res0 // scs
    spb = sd ** 3
    index = {tuple(p): i for i, p in enumerate(leaves)}
    lists = [[] for _ in range(spb * len(leaves))]
    idx = [[] for _ in range(spb * len(leaves))]
    r = np.float32(radius)
    for k, p in enumerate(w):
        if p[2] == np.float32(1.0e10) or np.any(p < 0):
            continue
        lo = [int(ctrunc_div(int(np.float32(p[a] - r)), scs)) * scs for a in range(3)]
        hi = [int(ctrunc_div(int(np.float32(p[a] + np.float32(1) + r)), scs)) * scs for a in range(3)]
        for x in range(lo[0], hi[0] + 1, scs):
            for y in range(lo[1], hi[1] + 1, scs):
                for z in range(lo[2], hi[2] + 1, scs):
                    b = (x // res0 * res0, y // res0 * res0, z // res0 * res0)
                    if b not in index:
                        continue
                    lx, ly, lz = (x - b[0]) // scs, (y - b[1]) // scs, (z - b[2]) // scs
                    lists[index[b] * spb + (lz * sd + ly) * sd + lx].append(p)
                    idx[index[b] * spb + (lz * sd + ly) * sd + lx].append(k)
    cnt = np.array([len(l) for l in lists], np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if len(cnt) else cnt
    pos = np.zeros((len(lists), 3), np.int64)
    nid = np.zeros(len(lists), np.int64)
    for i, b in enumerate(leaves):
        for s in range(spb):
            pos[i * spb + s] = (b[0] + s % sd * scs, b[1] + (s // sd) % sd * scs, b[2] + (s // sd // sd) * scs)
            nid[i * spb + s] = i
    pnt = np.asarray([p for l in lists for p in l], np.float32).reshape(-1, 3)
    return dict(cnt=cnt, off=off, pos=pos, nid=nid, pnt=pnt, idx=idx)


def density_grid(pts, radius, lo, hi):
    """FP64 max(0, max_p (r - |p - v|)) on integer voxels v in [lo, hi)."""
    lo = np.asarray(lo, np.int64)
    hi = np.asarray(hi, np.int64)
    g = np.zeros(tuple(hi - lo), np.float64)
    r = float(np.float32(radius))
    reach = int(np.ceil(r)) + 1
    for p in np.asarray(pts, np.float32).astype(np.float64):
        c = np.floor(p).astype(np.int64)
        a = np.maximum(c - reach, lo)
        b = np.minimum(c + reach + 1, hi)
        if np.any(a >= b):
            continue
        xs, ys, zs = (np.arange(a[k], b[k], dtype=np.float64) for k in range(3))
        jx = (p[0] - xs)[:, None, None]
        jy = (p[1] - ys)[None, :, None]
        jz = (p[2] - zs)[None, None, :]
        val = r - np.sqrt(jx * jx + jy * jy + jz * jz)
        sl = tuple(slice(a[k] - lo[k], b[k] - lo[k]) for k in range(3))
        np.maximum(g[sl], val, out=g[sl])
    return g


def brick_interiors(atlas, leaves, slots, w=8):
    """(B, w, w, w) interiors [x, y, z] of the bricks from a (Z, Y, X) atlas."""
    out = np.zeros((len(leaves), w, w, w), atlas.dtype)
    for i, s in enumerate(slots):
        blk = atlas[s[2]:s[2] + w, s[1]:s[1] + w, s[0]:s[0] + w]
        out[i] = blk.transpose(2, 1, 0)
    return out


def quantize16(w, mn, rg):
    """u16((w - mn) / rg * 65535) with one float32 rounding per operation,
    truncated; saturated to [0, 65535]."""
    w = np.asarray(w, np.float32)
    q = ((w - np.float32(mn)) / np.float32(rg)) * np.float32(65535)
    q = np.where(q >= np.float32(65535), np.float32(65535), np.where(q > 0, q, np.float32(0)))
    return np.trunc(q).astype(np.uint16)


def fma32(a, b, c):
    """Correctly rounded float32 a * b + c, from the exact rational value."""
    from fractions import Fraction
    exact = Fraction(float(np.float32(a))) * Fraction(float(np.float32(b))) + Fraction(float(np.float32(c)))
    x = np.float32(float(exact))
    best = None
    for cand in (np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - exact)
        if best is None or err < best[0] or (err == best[0] and int(cand.view(np.uint32)) % 2 == 0):
            best = (err, cand)
    return best[1]


def dequantize16(u, mn, rg):
    """u / 65535.0f * rg + mn, the quotient rounded to float32 and the
    product and sum fused (one rounding)."""
    return np.float32(fma32(np.float32(np.float32(u) / np.float32(65535)), rg, mn))


def incremental_model(pools, clists, pts, radius, D, accum, log2dims=(3, 3, 3, 4, 5)):
    """Expected pools after ActivateIncreBricksGPU (src/gvdb_volume_gvdb.cpp:
    1020-1156) from the pools before it, written with position dictionaries:

      found(lev, pos)  = a node at `lev` with corner `pos` is reachable from the
                         top node (the lowest level holding one node) through
                         child slots;
      keys             = every node corner of the per-axis box
                         [int(w) - r, int(w) + r] at each level below D that is
                         not found; found leaves of the boxes get marker 1;
      no key           -> nothing changes;
      otherwise        leaves take their markers (else 0, or their flags when
                       accumulating), new nodes are appended per level in
                       ascending (x, y, z) with flag 1 and empty child lists;
                       top-down, every flagged node of levels D-1..0 is put in
                       the slot of the reachable node one level up containing
                       it; unflagged leaves and then, per level 1..D-1, nodes
                       without children leave their parent's slot (parentless
                       nodes of those levels get flag 0; nodes with a child
                       get flag 1); the leaf used count is the flagged leaves.

    pools[lev]: list of dicts (pos, flags, parent, clist); clists[lev]: list of
    lists.  Returns (pools, clists, used0) as new objects."""
    import copy
    res, rng = tree_ranges(log2dims)
    P = copy.deepcopy(pools)
    C = copy.deepcopy(clists)
    nlev = len(P)

    def top_level():
        t = 1
        for n in range(nlev - 1, -1, -1):
            if len(P[n]) == 1:
                t = n
        return t

    def reachable():
        t = top_level()
        out = {t: {0}}
        for lev in range(t, 0, -1):
            out[lev - 1] = set()
            for i in out[lev]:
                cl = P[lev][i]["clist"]
                if cl == ID_UNDEFL:
                    continue
                for v in C[lev][cl >> 16]:
                    if v != ID_UNDEF64:
                        out[lev - 1].add(v >> 16)
        return out

    def lookup(reach, lev, pos):
        # getPosParent answers only below the top node's level
        if lev >= top_level():
            return None
        for i in reach.get(lev, ()):
            if tuple(P[lev][i]["pos"]) == tuple(pos):
                return i
        return None

    def containing(reach, lev, pos):
        """getParent's answer: the top node itself at the top level, else the
        reachable node at `lev` whose box holds pos (None if none)."""
        t = top_level()
        if lev == t:
            return 0
        if lev > t:
            return None
        for i in reach.get(lev, ()):
            c = P[lev][i]["pos"]
            if all(c[a] <= pos[a] < c[a] + rng[lev] for a in range(3)):
                return i
        return None

    reach = reachable()
    nleaf = len(P[0])
    if nleaf == 0:
        return P, C, None
    marker = [P[0][i]["flags"] if accum else 0 for i in range(nleaf)]
    keys = set()
    for p in np.asarray(pts, np.float32):
        if np.any(p < 0):
            continue
        iw = [int(p[a]) for a in range(3)]
        lo = [int(np.float32(np.float32(iw[a]) - np.float32(radius))) for a in range(3)]
        hi = [int(np.float32(np.float32(iw[a]) + np.float32(radius))) for a in range(3)]
        for lev in range(D):
            rr = rng[lev]
            ax = [range(int(ctrunc_div(lo[a], rr)) * rr, int(ctrunc_div(hi[a], rr)) * rr + 1, rr) for a in range(3)]
            for x in ax[0]:
                for y in ax[1]:
                    for z in ax[2]:
                        i = lookup(reach, lev, (x, y, z))
                        if i is None:
                            keys.add((lev, x // rr, y // rr, z // rr))
                        elif lev == 0:
                            marker[i] = 1
    if not keys:
        return P, C, None
    for i in range(nleaf):
        P[0][i]["flags"] = marker[i]
    for (lev, x, y, z) in sorted(keys):
        rr = rng[lev]
        nd = dict(pos=(x * rr, y * rr, z * rr), flags=1, parent=ID_UNDEFL, clist=ID_UNDEFL, value=(-1, -1, -1))
        if lev > 0:
            nd["clist"] = elem(1, lev, len(C[lev]))
            C[lev].append([ID_UNDEF64] * res[lev] ** 3)
        P[lev].append(nd)

    def slot(lev_p, pnode, pos):
        c = [(pos[a] - pnode["pos"][a]) * res[lev_p] // rng[lev_p] for a in range(3)]
        return c, (c[2] * res[lev_p] + c[1]) * res[lev_p] + c[0]

    for lev in range(D - 1, -1, -1):
        reach = reachable()
        for i, nd in enumerate(P[lev]):
            if not nd["flags"]:
                continue
            pid = containing(reach, lev + 1, nd["pos"])
            if pid is None:
                continue
            c, b = slot(lev + 1, P[lev + 1][pid], nd["pos"])
            if any(v < 0 or v >= res[lev + 1] for v in c):
                continue
            nd["parent"] = elem(0, lev + 1, pid)
            C[lev + 1][P[lev + 1][pid]["clist"] >> 16][b] = elem(0, lev, i)

    def delink(lev, nd):
        pid = containing(reachable(), lev + 1, nd["pos"])
        _, b = slot(lev + 1, P[lev + 1][pid], nd["pos"])
        nd["parent"] = ID_UNDEF64
        C[lev + 1][P[lev + 1][pid]["clist"] >> 16][b] = ID_UNDEF64

    for nd in P[0]:
        if not nd["flags"] and nd["parent"] != ID_UNDEF64:
            delink(0, nd)
    for lev in range(1, D):
        for i, nd in enumerate(P[lev]):
            if nd["parent"] == ID_UNDEF64:
                nd["flags"] = 0
                continue
            if any(v != ID_UNDEF64 for v in C[lev][i]):
                nd["flags"] = 1
                continue
            nd["flags"] = 0
            delink(lev, nd)
    used0 = sum(1 for nd in P[0] if nd["flags"])
    return P, C, used0


def pools_as_dicts(o, nlev=5):
    """Node pools and child lists of an Oracle (or a GPU volume) as the
    lists incremental_model takes."""
    P, C = [], []
    for lev in range(nlev):
        a = o.pool(0, lev)
        P.append([dict(pos=tuple(int(v) for v in n["mPos"]), flags=int(n["mFlags"]), parent=int(n["mParent"]),
                       clist=int(n["mChildList"]), value=tuple(int(v) for v in n["mValue"])) for n in a])
        cl = o.pool(1, lev) if lev > 0 else np.zeros((0, 0), np.uint64)
        C.append([[int(v) for v in row] for row in cl])
    return P, C
