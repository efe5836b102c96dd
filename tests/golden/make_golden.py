"""Writes the golden fixtures under tests/golden/ from the CPU oracle (FP64).

The reference ships no golden vectors for this path, so these fixtures are
the oracle's output on small seeded inputs, after the oracle itself has been
pinned by tests/test_oracle_cpu.py (brute-force models and closed forms).
They let the GPU tests check the HIP path against stored numbers as well as
against a live oracle run.

    python tests/golden/make_golden.py        # rewrites the .npz files

Each fixture holds: the input points, the radius and subcell size, the node
pools (positions, parents, child-list references, child lists), atlas slot
corners, subcell counts, the padded 10^3 bricks of the gathered density after
UpdateApron (FP64 rounded to float32, in pool order, [brick, z, y, x]), and two
64x64 float RGBA renders (emission-absorption, isosurface) with their scene
blocks and transfer functions.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))

import oracle as O                      # noqa: E402
from gvdb_amd import workloads as W     # noqa: E402

FIXTURES = {
    "sphere_r2": dict(kind="sphere", n=1000, seed=101, radius=2.0, scs=4, centre=(200.0, 180.0, 220.0), shell=20.0),
    "cloud_r1p5": dict(kind="cloud", n=600, seed=102, radius=1.5, scs=4),
}


def points(spec):
    if spec["kind"] == "sphere":
        return W.sphere_shell(spec["n"], seed=spec["seed"], radius=spec["shell"], centre=spec["centre"])
    g = np.random.Generator(np.random.PCG64(spec["seed"]))
    return (100.0 + 60.0 * g.random((spec["n"], 3))).astype(np.float32)


def scenes(centre):
    ea = W.ea_scene(64, 64, centre=centre, dist=160.0)
    iso = W.iso_scene(64, 64, centre=centre, dist=160.0)
    return ea, iso


def build(name, spec):
    pts = points(spec)
    r, scs = spec["radius"], spec["scs"]
    o = O.Oracle()
    assert o.rebuild_topology(pts, r) == 0 and o.undefined_events() == 0
    o.finish_topology()
    assert o.update_atlas() == 0
    o.clear_channel()
    rc, n = o.insert_points_subcell(pts, r, scs)
    assert rc == 0
    o.gather_density(r, scs, 64)
    o.update_apron(0.0, 64)
    out = dict(points=pts, radius=np.float32(r), scs=np.int32(scs), entries=np.int64(n))
    for lev in range(5):
        p = o.pool(0, lev)
        out["pos%d" % lev] = p["mPos"]
        out["parent%d" % lev] = p["mParent"]
        out["childref%d" % lev] = p["mChildList"]
        if lev > 0:
            out["clist%d" % lev] = o.pool(1, lev)
    leaves = o.pool(0, 0)
    out["slots"] = leaves["mValue"]
    out["sc_cnt"] = o.subcells(n)["cnt"]
    a = o.atlas(64)
    c = leaves["mValue"] - 1
    out["bricks"] = np.stack([a[z:z + 10, y:y + 10, x:x + 10] for x, y, z in c]).astype(np.float32)
    centre = tuple(float(v) for v in pts.mean(0))
    ea, iso = scenes(centre)
    for key, sc, shading in (("ea", ea, 7), ("iso", iso, 4)):
        blk = O.scene_block(sc)
        tf = O.linear_tf(*sc["tf"])
        out["scene_" + key] = blk
        out["img_" + key] = o.render(blk, tf, shading, 64)
    meta = dict(spec, entries=int(n), bricks=int(len(leaves)), centre=centre, scenes=dict(ea=ea, iso=iso))
    return out, meta


def main():
    for name, spec in FIXTURES.items():
        out, meta = build(name, spec)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(meta, f, indent=1, default=list)
        print(name, meta["bricks"], "bricks", meta["entries"], "entries")


if __name__ == "__main__":
    main()
