"""Seeded synthetic point sets and scene settings for the point-to-volume path.

Every generator uses numpy's PCG64 bit generator (deterministic across
platforms for a given seed) and returns an (N, 3) float32 array of index-space
positions.  Points are generated in float64 and cast to float32 once.

Workloads (tree log2 dims {3,3,3,4,5} leaf first, 8^3 bricks, apron 1):
  sphere_shell   262,144 points, radius 120 + N(0, 1.5), centre (256,256,256)
  slab_plane     4,194,304 points uniform in [56,456]x[20,220]x[56,456] plus
                 200,000 points uniform on the plane y = 10, x,z in [56,456]
  noise_cloud    points uniform in the cells (8 voxels wide in a 512^3 box)
                 where 3-octave value noise exceeds a threshold
  clusters       spherical Gaussian clusters (sigma 20-80) with centres in a
                 2048^3 index space; points outside [0, 2048) are dropped
  batch_grids    64 grids, point counts log-uniform in [10k, 2M], one of the
                 three shapes above each (the batched workload)
  serve_grids    up to 200 grids, point counts log-uniform in [50k, 4M]
                 (the serving workload's resident grids)
"""
import numpy as np

TREE_LOG2 = (3, 3, 3, 4, 5)      # leaf first, i.e. Configure(5, 4, 3, 3, 3)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def sphere_shell(n=262144, seed=1, radius=120.0, centre=(256.0, 256.0, 256.0), sigma=1.5):
    g = _rng(seed)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius + sigma * g.standard_normal(n)
    p = np.asarray(centre, dtype=np.float64)[None, :] + d * r[:, None]
    return p.astype(np.float32)


def slab_plane(n_slab=4194304, n_plane=200000, seed=2, lo=(56.0, 20.0, 56.0), hi=(456.0, 220.0, 456.0), plane_y=10.0):
    g = _rng(seed)
    lo = np.asarray(lo)
    hi = np.asarray(hi)
    s = lo + (hi - lo) * g.random((n_slab, 3))
    q = np.empty((n_plane, 3))
    q[:, 0] = lo[0] + (hi[0] - lo[0]) * g.random(n_plane)
    q[:, 1] = plane_y
    q[:, 2] = lo[2] + (hi[2] - lo[2]) * g.random(n_plane)
    return np.concatenate([s, q]).astype(np.float32)


def _value_noise(p, g, octaves=3, base_cells=8):
    """3-octave trilinear value noise in [0, 1] at points p in [0, 1)^3."""
    total = np.zeros(len(p))
    amp_sum = 0.0
    for o in range(octaves):
        cells = base_cells << o
        lattice = g.random((cells + 1, cells + 1, cells + 1))
        q = p * cells
        i = np.minimum(np.floor(q).astype(np.int64), cells - 1)
        f = q - i
        f = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            wx = f[:, 0] if dx else 1 - f[:, 0]
            for dy in (0, 1):
                wy = f[:, 1] if dy else 1 - f[:, 1]
                for dz in (0, 1):
                    wz = f[:, 2] if dz else 1 - f[:, 2]
                    acc += wx * wy * wz * lattice[i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz]
        amp = 0.5 ** o
        total += amp * acc
        amp_sum += amp
    return total / amp_sum


def noise_cloud(n, seed=3, box=512.0, thresh=0.55, margin=8.0, cells=64):
    """Points uniform inside the cells of a cells^3 grid over the box whose
    3-octave value noise (evaluated at the cell centres) exceeds `thresh`."""
    g = _rng(seed)
    c = (np.arange(cells) + 0.5) / cells
    cx, cy, cz = np.meshgrid(c, c, c, indexing="ij")
    centres = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1)
    keep = np.nonzero(_value_noise(centres, _rng(seed + 1000)) > thresh)[0]
    if len(keep) == 0:
        keep = np.arange(len(centres))
    pick = keep[g.integers(0, len(keep), n)]
    ijk = np.stack(np.unravel_index(pick, (cells, cells, cells)), 1).astype(np.float64)
    p = (ijk + g.random((n, 3))) / cells
    return (margin + p * (box - 2 * margin)).astype(np.float32)


def clusters(n=16777216, k=32, seed=4, space=2048.0, sig=(20.0, 80.0)):
    g = _rng(seed)
    centres = g.random((k, 3)) * space
    sigmas = sig[0] + (sig[1] - sig[0]) * g.random(k)
    lab = g.integers(0, k, n)
    p = centres[lab] + sigmas[lab, None] * g.standard_normal((n, 3))
    keep = np.all((p >= 0) & (p < space), axis=1)
    return p[keep].astype(np.float32)


def batch_grids(count=64, seed=3, n_lo=10000, n_hi=2000000):
    """The batched workload: `count` grids, log-uniform point counts, one of
    three shapes per grid.  Returns a list of (shape_name, points)."""
    g = _rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(np.log(n_lo) + (np.log(n_hi) - np.log(n_lo)) * g.random())))
        kind = int(g.integers(0, 3))
        s = seed * 1000 + i
        if kind == 0:
            out.append(("shell", sphere_shell(n, seed=s)))
        elif kind == 1:
            frac = 0.95
            out.append(("slab", slab_plane(int(n * frac), n - int(n * frac), seed=s)))
        else:
            out.append(("cloud", noise_cloud(n, seed=s)))
    return out


def serve_grids(count=200, seed=5, n_lo=50000, n_hi=4000000):
    """The serving workload's resident grids: `count` grids with initial
    point counts log-uniform in [n_lo, n_hi], one of the three shapes each
    (the batched workload's generator with the serving sizes)."""
    return batch_grids(count=count, seed=seed, n_lo=n_lo, n_hi=n_hi)


# ----------------------------------------------------------------- scenes --
def ea_scene(w=256, h=256, centre=(256.0, 256.0, 256.0), dist=600.0):
    """Emission-absorption render settings (shading SHADE_VOLUME = 7).

    Density scale 0.05: the transfer function maps density [0, 2] linearly to
    alpha [0, 1] and the extinction coefficient is 0.1, so the absorption per
    voxel is 0.05 * density."""
    return dict(
        width=w, height=h, shading=7,
        cam=dict(angs=(30.0, 20.0, 0.0), to=centre, dist=dist, dolly=1.0, fov=40.0, near=0.1, far=5000.0),
        light=(centre[0] + 300.0, centre[1] + 500.0, centre[2] + 200.0),
        backclr=(0.1, 0.1, 0.15, 1.0),
        extinct=(-0.1, 1.0, 0.0), steps=(0.5, 16.0, 0.1), cutoff=(0.005, 0.01, 0.0),
        vrange=(0.0, 0.0, 2.0), shadow=(0.8, 1.0, 0.0),
        tf=(0.0, 1.0, (1.0, 1.0, 1.0, 0.0), (1.0, 0.9, 0.7, 1.0)),
    )


def iso_scene(w=1024, h=1024, centre=(256.0, 120.0, 256.0), dist=900.0):
    """Isosurface settings (shading SHADE_TRILINEAR = 4): threshold 0.5,
    step 0.25, Phong with a shadow ray (shadow amount 0.8)."""
    return dict(
        width=w, height=h, shading=4,
        cam=dict(angs=(35.0, 30.0, 0.0), to=centre, dist=dist, dolly=1.0, fov=40.0, near=0.1, far=5000.0),
        light=(centre[0] + 400.0, centre[1] + 600.0, centre[2] + 300.0),
        backclr=(0.1, 0.1, 0.15, 1.0),
        extinct=(-1.1, 1.5, 0.0), steps=(0.25, 16.0, 0.1), cutoff=(0.005, 0.01, 0.0),
        vrange=(0.5, 0.0, 1.0), shadow=(0.8, 1.0, 0.0),
        tf=(0.0, 1.0, (0.0, 0.0, 0.0, 0.0), (1.0, 1.0, 1.0, 0.1)),
    )
