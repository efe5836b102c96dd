"""VBX file writer and reader restated from the reference (test infrastructure
only; nothing in the product imports it).

The byte layout follows SaveVBX / LoadVBX (source/gvdb_library/src/
gvdb_volume_gvdb.cpp:1626-1768 and :507-685) and GVDB_FILESPEC.txt:
little-endian, no padding, version 1.11 (:31-32), one grid, GVDB topology
in atlas layout.  Where the specification text and SaveVBX disagree (SaveVBX
writes the channel count before the atlas size and writes pool-1 counts per
level), SaveVBX is followed, since LoadVBX reads what it writes.
"""
import struct

import numpy as np

HEADER = struct.Struct("<BB12fi")              # major, minor, 4 x vec3f, grid count
GRID = struct.Struct("<256sBBB3fi3iiiQBiB3i3i")  # grid header up to the atlas resolution
TOPO = struct.Struct("<iQ")                    # levels, root
LEVEL = struct.Struct("<ii3iiiii")             # log2dim, res, range, cnt0, width0, cnt1, width1


def write_vbx(meta, pools0, pools1, atlases, transform=None):
    """meta: dict(levels, root, log2dims, res, ranges, apron, brick, axiscnt,
    bricks); pools0/pools1: per level raw bytes (rows x width); atlases: list
    of (type, numpy 3D array z,y,x).  Returns the file as bytes."""
    t = transform if transform is not None else [0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0]
    out = bytearray(HEADER.pack(1, 11, *t, 1))
    table = len(out)
    out += struct.pack("<Q", 0)
    grid_off = len(out)
    br = meta["brick"] + 2 * meta["apron"]
    ax = meta["axiscnt"] if atlases else (0, 0, 0)
    res3 = tuple(a * br for a in ax)
    asz = int(np.prod(res3)) * atlases[0][1].dtype.itemsize if atlases else 0
    out += GRID.pack(b"\0" * 256, ord("f"), 1, 0, 1.0, 1.0, 1.0, meta["bricks"], meta["brick"], meta["brick"],
                     meta["brick"], meta["apron"], len(atlases), asz, 2, 0, 0, *ax, *res3)
    out += TOPO.pack(meta["levels"], meta["root"])
    for n in range(meta["levels"]):
        w0 = 64
        w1 = 0 if n == 0 else 8 * meta["res"][n] ** 3
        out += LEVEL.pack(meta["log2dims"][n], meta["res"][n], meta["ranges"][n], meta["ranges"][n],
                          meta["ranges"][n], len(pools0[n]) // w0, w0, (len(pools1[n]) // w1) if w1 else 0, w1)
    for p in pools0:
        out += p
    for n, p in enumerate(pools1):
        if n:
            out += p
    for typ, a in atlases:
        out += struct.pack("<ii", typ, a.dtype.itemsize)
        out += np.ascontiguousarray(a).tobytes()
    struct.pack_into("<Q", out, table, grid_off)
    return bytes(out)


def read_vbx(buf):
    """Parses a VBX file written by SaveVBX (one grid, no bitmasks) into a dict
    with the header fields, per-level pool bytes and the atlases (z, y, x)."""
    off = 0
    h = HEADER.unpack_from(buf, off)
    off += HEADER.size
    major, minor, xf, ngrids = h[0], h[1], h[2:14], h[14]
    offs = struct.unpack_from("<%dQ" % ngrids, buf, off)
    off += 8 * ngrids
    g = GRID.unpack_from(buf, off)
    off += GRID.size
    (name, dtype, comps, compress, vs0, vs1, vs2, bricks, bx, by, bz, apron, nchan, asz, topo, reuse, layout,
     a0, a1, a2, r0, r1, r2) = g
    levels, root = TOPO.unpack_from(buf, off)
    off += TOPO.size
    lev = []
    for n in range(levels):
        lev.append(LEVEL.unpack_from(buf, off))
        off += LEVEL.size
    pools0, pools1 = [], []
    for n in range(levels):
        cnt, w = lev[n][5], lev[n][6]
        pools0.append(bytes(buf[off:off + cnt * w]))
        off += cnt * w
    for n in range(levels):
        cnt, w = lev[n][7], lev[n][8]
        pools1.append(bytes(buf[off:off + cnt * w]))
        off += cnt * w
    atlases = []
    for c in range(nchan):
        typ, stride = struct.unpack_from("<ii", buf, off)
        off += 8
        dt = {4: np.float32, 2: np.float16}[stride]
        n = r0 * r1 * r2
        a = np.frombuffer(buf, dt, n, off).reshape(r2, r1, r0)
        off += n * stride
        atlases.append((typ, a))
    return dict(major=major, minor=minor, transform=xf, grids=ngrids, offsets=offs, name=name, dtype=chr(dtype),
                components=comps, compression=compress, voxelsize=(vs0, vs1, vs2), bricks=bricks,
                brick=(bx, by, bz), apron=apron, channels=nchan, atlas_bytes=asz, topology=topo, reuse=reuse,
                layout=layout, axiscnt=(a0, a1, a2), axisres=(r0, r1, r2), levels=levels, root=root,
                level_info=lev, pools0=pools0, pools1=pools1, atlases=atlases, end=off)
