"""Precision report for the optional FP16 atlas channel (run on a GPU host).

Workload: 16,777,216 points drawn from 32 spherical Gaussian clusters
(sigma 20-80 voxels, centres uniform in a 2048^3 index space; points outside
[0, 2048) dropped), radius 2, subcell 4, tree {3,3,3,4,5}.  The density atlas
is gathered on the GPU into an FP32 channel and into an FP16 channel (FP32
arithmetic, FP16 storage); both are compared voxel by voxel with the CPU
oracle's FP64 gather, and a 2048x2048 emission-absorption render from each
(FP32 sampling) is compared pixel by pixel with the oracle's FP32 march (the
reference's arithmetic) and its FP64 march.

    python tests/precision_report.py [--points N] [--width W] [--out FILE]

Prints a markdown table (and writes it to --out).  The oracle here is the
checker, not the thing measured.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import helpers as H   # noqa: E402


def log(msg):
    print("[%7.1fs] %s" % (time.perf_counter() - T0, msg), flush=True)


def errors(a, b, mask):
    a = a[mask].astype(np.float64)
    b = b[mask]
    d = np.abs(a - b)
    nz = b > 1e-4
    rel = d[nz] / b[nz]
    return dict(max_abs=float(d.max()), rms=float(np.sqrt(np.mean(d * d))), max_rel=float(rel.max()) if nz.any() else 0.0,
                p99_rel=float(np.percentile(rel, 99)) if nz.any() else 0.0,
                zero_mismatch=int(np.count_nonzero((b == 0) & (a != 0))), voxels=int(a.size))


def main():
    global T0
    T0 = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=16777216)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    import gvdb_amd as G
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.clusters(args.points, k=32, seed=4, space=2048.0, sig=(20.0, 80.0))
    log("%d points" % len(pts))
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    sc = W.ea_scene(args.width, args.width, centre=tuple((lo + hi) / 2), dist=float(1.6 * np.max(hi - lo)))
    res = {}
    for name, t in (("fp32", G.T_FLOAT), ("fp16", G.T_HALF)):
        v, n = H.gpu_volume(G, pts, 2.0, atlas_type=t)
        v.apply_scene(sc)
        v.AddRenderBuf(0, args.width, args.width, 16)
        v.Render(G.SHADE_VOLUME, 0, 0)
        img = v.ReadRenderBuf(0, args.width, args.width, 16)
        atlas = v.atlas(0, np.float32 if t == G.T_FLOAT else np.float16)
        res[name] = dict(atlas=atlas, img=img, bricks=len(v.pool(0, 0)), entries=n, blk=v.scene_block(args.width, args.width))
        v.close()
        log("GPU %s: %d bricks, %d entries" % (name, res[name]["bricks"], n))
    O.set_threads(os.cpu_count() or 1)
    o = O.Oracle()
    assert o.rebuild_topology(pts, 2.0) == 0
    o.finish_topology()
    assert o.update_atlas() == 0
    o.clear_channel()
    log("oracle topology + atlas")
    rc, m = o.insert_points_subcell(pts, 2.0, 4)
    assert rc == 0 and m == res["fp32"]["entries"]
    log("oracle insert: %d entries" % m)
    o.gather_density(2.0, 4, 64)
    o.update_apron(0.0, 64)
    log("oracle FP64 gather + apron")
    ref = o.atlas(64)
    mask = H.used_brick_mask(ref.shape, o.atlas_map(), o.atlas_info()["cnt"])
    e32 = errors(res["fp32"]["atlas"], ref, mask)
    e16 = errors(res["fp16"]["atlas"], ref, mask)
    del ref
    log("voxel errors done")
    tf = O.linear_tf(*sc["tf"])
    rimg, capped = o.render(res["fp32"]["blk"], tf, 7, 64, capped=True)
    log("oracle FP64 march")
    o.gather_density(2.0, 4, 32)
    o.update_apron(0.0, 32)
    fimg, fcap = o.render(res["fp32"]["blk"], tf, 7, 32, capped=True)
    log("oracle FP32 gather + march")
    whole = ~(capped | fcap)
    # rays on which the FP32 and FP64 walks agree: the conditioned part of the image
    agree = whole & (np.abs(fimg.astype(np.float64) - rimg).max(-1) <= 1e-3)
    lines = ["# FP16 vs FP32 atlas against the FP64 oracle: 32 Gaussian clusters",
             "",
             "%d points (32 clusters, sigma 20-80, centres in 2048^3), radius 2, subcell 4; %d bricks, %d subcell "
             "entries; %dx%d emission-absorption render." % (len(pts), res["fp32"]["bricks"], m, args.width, args.width),
             "",
             "Command (repository root, one MI355X): `python tests/precision_report.py --out <file>`",
             "",
             "## Atlas densities against the FP64 gather",
             "",
             "| atlas | voxels | max abs | RMS abs | max rel (ρ>1e-4) | p99 rel | zero mismatches |",
             "|---|---|---|---|---|---|---|"]
    for name, e in (("FP32", e32), ("FP16", e16)):
        lines.append("| %s | %d | %.3g | %.3g | %.3g | %.3g | %d |" % (name, e["voxels"], e["max_abs"], e["rms"],
                                                                   e["max_rel"], e["p99_rel"], e["zero_mismatch"]))
    lines += ["",
              "## Images",
              "",
              "The march is the reference's FP32 algorithm (FP32 DDA with its 1e-3 epsilon, MAX_ITER = 256 walk "
              "cap).  Run in FP64 on the same scene it is not the same walk: at coordinates near 4000 the epsilon is "
              "four float ulps, so a child node's exit time can equal the parent's limit in FP32 and the walk then "
              "leaves the loop, where FP64 pops to the parent and goes on.  The GPU march performs the FP32 "
              "operations one by one (no contraction), so it takes the FP32 walk.  The images are compared with "
              "both the oracle's FP32 march (the reference's arithmetic: the parity claim) and its FP64 march (the "
              "error against FP64 asked of each result); the FP32-vs-FP64 row of the oracle itself shows how much "
              "of the latter is the algorithm's own.",
              "",
              "%d of %d rays (%.1f%%) stop at the walk cap in the FP64 or FP32 march; where they stop depends on the "
              "exact step sequence, so they are listed apart.  %d rays (%.1f%%) are within the cap and end within "
              "1e-3 of each other in the FP32 and FP64 marches: the well-conditioned rays." % (
                  int((~whole).sum()), whole.size, 100.0 * (~whole).mean(), int(agree.sum()), 100.0 * agree.mean()),
              "", "| image | against | rays | pixels | max abs | RMSE |", "|---|---|---|---|---|---|"]
    pairs = (("GPU, FP32 atlas", res["fp32"]["img"], "oracle FP32 march", fimg),
             ("GPU, FP16 atlas", res["fp16"]["img"], "oracle FP32 march", fimg),
             ("GPU, FP32 atlas", res["fp32"]["img"], "oracle FP64 march", rimg),
             ("GPU, FP16 atlas", res["fp16"]["img"], "oracle FP64 march", rimg),
             ("GPU, FP16 atlas", res["fp16"]["img"], "GPU, FP32 atlas", res["fp32"]["img"]),
             ("oracle FP32 march", fimg, "oracle FP64 march", rimg))
    sels = (("well-conditioned", agree), ("within the cap", whole), ("all", np.ones_like(whole)))
    for name, a, what, b in pairs:
        p = np.abs(a.astype(np.float64) - b.astype(np.float64))
        for sel, m in sels:
            q = p[m]
            lines.append("| %s | %s | %s | %d | %.3g | %.3g |" % (name, what, sel, m.sum(),
                                                              float(q.max()) if q.size else 0.0,
                                                              float(np.sqrt(np.mean(q * q))) if q.size else 0.0))
    text = "\n".join(lines) + "\n"
    print(text, flush=True)
    if args.out:
        open(args.out, "w").write(text)


if __name__ == "__main__":
    main()
