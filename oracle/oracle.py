"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.  See gvdb_oracle.cpp for the reference
file:line each routine restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

SCN_FLOATS = 68
NODE_DTYPE = np.dtype([
    ("mLev", "u1"), ("mFlags", "u1"), ("mPriority", "u1"), ("pad", "u1"),
    ("mPos", "<i4", 3), ("mValue", "<i4", 3), ("mVRange", "<f4", 3),
    ("mParent", "<u8"), ("mChildList", "<u8"), ("mMask", "<u8")])
ATLASNODE_DTYPE = np.dtype([("mPos", "<i4", 3), ("mLeafNode", "<i4")])
TF_SIZE = 16384

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i, f, fp, ip = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p
        sig = {
            "orc_new": (vp, [i, ip]), "orc_free": (None, [vp]),
            "orc_set_channel_default": (None, [vp, i, i, i]), "orc_add_channel": (None, [vp, i]),
            "orc_set_epsilon": (None, [vp, f]), "orc_set_fp64": (None, [vp, i]), "orc_undefined_events": (i, [vp]),
            "orc_rebuild_topology": (i, [vp, fp, i, f, fp]), "orc_rebuild_topology_cpu": (i, [vp, fp, i, fp]),
            "orc_accumulate_topology": (i, [vp, fp, i, f, fp, i]), "orc_request_full_rebuild": (None, [vp, i]),
            "orc_pool_used": (ctypes.c_ulonglong, [vp, i, i]),
            "orc_clear": (None, [vp]), "orc_finish_topology": (None, [vp]), "orc_update_atlas": (i, [vp]),
            "orc_clear_channel": (None, [vp]), "orc_fill_channel": (None, [vp, f]),
            "orc_pool_count": (ctypes.c_ulonglong, [vp, i, i]), "orc_pool_stride": (ctypes.c_ulonglong, [vp, i, i]),
            "orc_pool_copy": (None, [vp, i, i, vp]), "orc_top_lev": (i, [vp]), "orc_bounds": (None, [vp, fp]),
            "orc_atlas_info": (None, [vp, ip]), "orc_atlas_copy": (None, [vp, i, vp]),
            "orc_atlas_map_copy": (None, [vp, vp]),
            "orc_insert_points_subcell": (i, [vp, i, fp, i, f, fp, ip]), "orc_num_subcells": (i, [vp]),
            "orc_subcell_copy": (None, [vp, i, vp]),
            "orc_gather_density": (i, [vp, i, f, i, i]), "orc_update_apron": (i, [vp, f, i]),
            "orc_gather_level_set": (i, [vp, i, f, i, i]),
            "orc_insert_points_subcell_fp16": (i, [vp, i, fp, i, f, fp, ip]),
            "orc_subcell16_copy": (None, [vp, vp]),
            "orc_gather_level_set_fp16": (i, [vp, i, f, i]),
            "orc_compute": (i, [vp, i, i, f, f, f, i, i, f, i]), "orc_random3": (f, [f, f, f]),
            "orc_atlas_set": (None, [vp, i, vp]), "orc_hash": (ctypes.c_uint, [ctypes.c_uint]),
            "orc_camera_scn": (None, [fp, i, i, fp]), "orc_scn_floats": (i, []),
            "orc_linear_tf": (None, [f, f, fp, fp, fp]),
            "orc_render": (i, [vp, fp, fp, i, i, fp, vp]), "orc_stage_ms": (ctypes.c_double, [vp, i]),
            "orc_max_threads": (i, []), "orc_set_threads": (None, [i]),
            "orc_add_color_channel": (None, [vp]), "orc_set_color_channel": (None, [vp, i]),
            "orc_clear_color": (None, [vp]), "orc_color_copy": (None, [vp, vp]), "orc_color_set": (None, [vp, vp]),
            "orc_insert_points_subcell_clr": (i, [vp, i, fp, vp, i, f, fp, ip]),
            "orc_gather_color": (i, [vp, i, i, i]), "orc_update_apron_color": (i, [vp, f]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        assert L.orc_scn_floats() == SCN_FLOATS
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


class Oracle:
    """One GVDB volume restated on the CPU (Configure + AddChannel done)."""

    def __init__(self, log2dims=(3, 3, 3, 4, 5), apron=1, axiscnt=(16, 16, 1), fp64=True):
        L = lib()
        dims = np.asarray(log2dims, dtype=np.int32)
        self.levels = len(dims)
        self.h = L.orc_new(self.levels, _p(dims))
        L.orc_set_fp64(self.h, int(fp64))
        L.orc_set_channel_default(self.h, *axiscnt)
        L.orc_add_channel(self.h, apron)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    # topology ------------------------------------------------------------
    def rebuild_topology(self, pts, radius, orig=(0, 0, 0)):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        return lib().orc_rebuild_topology(self.h, _p(pts), len(pts), float(radius), _p(f3(orig)))

    def accumulate_topology(self, pts, radius, orig=(0, 0, 0), depth=1):
        """AccumulateTopology: the incremental build keeping existing leaves."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        return lib().orc_accumulate_topology(self.h, _p(pts), len(pts), float(radius), _p(f3(orig)), int(depth))

    def request_full_rebuild(self, on=True):
        lib().orc_request_full_rebuild(self.h, int(on))

    def pool_used(self, grp, lev):
        return int(lib().orc_pool_used(self.h, grp, lev))

    def rebuild_topology_cpu(self, pts, orig=(0, 0, 0)):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        return lib().orc_rebuild_topology_cpu(self.h, _p(pts), len(pts), _p(f3(orig)))

    def clear(self):
        lib().orc_clear(self.h)

    def finish_topology(self):
        lib().orc_finish_topology(self.h)

    def update_atlas(self):
        return lib().orc_update_atlas(self.h)

    def clear_channel(self):
        lib().orc_clear_channel(self.h)

    def fill_channel(self, value):
        """Sets every atlas voxel to `value` (uniform field for closed-form checks)."""
        lib().orc_fill_channel(self.h, float(value))

    def undefined_events(self):
        return lib().orc_undefined_events(self.h)

    def pool(self, grp, lev):
        L = lib()
        n = L.orc_pool_count(self.h, grp, lev)
        s = L.orc_pool_stride(self.h, grp, lev)
        buf = np.zeros(int(n * s), dtype=np.uint8)
        if n:
            L.orc_pool_copy(self.h, grp, lev, _p(buf))
        if grp == 0:
            return buf.view(NODE_DTYPE)
        return buf.view(np.uint64).reshape(int(n), -1) if n else np.zeros((0, 0), np.uint64)

    def top_lev(self):
        return lib().orc_top_lev(self.h)

    def bounds(self):
        b = np.zeros(12, np.float32)
        lib().orc_bounds(self.h, _p(b))
        return dict(obj_min=b[0:3], obj_max=b[3:6], pos_min=b[6:9], pos_max=b[9:12])

    def atlas_info(self):
        b = np.zeros(7, np.int32)
        lib().orc_atlas_info(self.h, _p(b))
        return dict(cnt=tuple(b[0:3]), res=tuple(b[3:6]), used=int(b[6]))

    def atlas(self, precision=32):
        info = self.atlas_info()
        rx, ry, rz = info["res"]
        a = np.zeros((rz, ry, rx), np.float64 if precision == 64 else np.float32)
        lib().orc_atlas_copy(self.h, precision, _p(a))
        return a

    def atlas_map(self):
        info = self.atlas_info()
        n = int(np.prod(info["cnt"]))
        a = np.zeros(n, ATLASNODE_DTYPE)
        lib().orc_atlas_map_copy(self.h, _p(a))
        return a

    # points ---------------------------------------------------------------
    def insert_points_subcell(self, pts, radius, scs=4, trans=(0, 0, 0), colors=None):
        """InsertPointsSubcell; with `colors` (a uint32 per point) the entries
        carry them, as with SetPoints' colour buffer."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        n = ctypes.c_int(0)
        if colors is None:
            rc = lib().orc_insert_points_subcell(self.h, scs, _p(pts), len(pts), float(radius), _p(f3(trans)),
                                                 ctypes.byref(n))
        else:
            c = np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
            assert len(c) >= len(pts)
            rc = lib().orc_insert_points_subcell_clr(self.h, scs, _p(pts), _p(c), len(pts), float(radius),
                                                     _p(f3(trans)), ctypes.byref(n))
        return rc, n.value

    # the colour channel (one uchar4 channel on the same atlas) -----------
    def add_color_channel(self):
        lib().orc_add_color_channel(self.h)

    def set_color_channel(self, on=True):
        """SetColorChannel: the march reads the colour channel."""
        lib().orc_set_color_channel(self.h, int(on))

    def clear_color(self):
        lib().orc_clear_color(self.h)

    def color_atlas(self):
        """The colour channel in the atlas layout, a uint32 (uchar4) per voxel."""
        rx, ry, rz = self.atlas_info()["res"]
        a = np.zeros((rz, ry, rx), np.uint32)
        lib().orc_color_copy(self.h, _p(a))
        return a

    def set_color_atlas(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        assert a.size == np.prod(self.atlas_info()["res"])
        lib().orc_color_set(self.h, _p(a))

    def gather_color(self, scs=4, accumulate=False, exact=True):
        """GatherDensity's colour channel from the entries' colours; exact:
        the mean in integers, else the reference's float sum in list order."""
        return lib().orc_gather_color(self.h, scs, int(accumulate), int(exact))

    def update_apron_color(self, boundval=0.0):
        return lib().orc_update_apron_color(self.h, float(boundval))

    def subcells(self, length):
        L = lib()
        s = L.orc_num_subcells(self.h)
        out = {}
        for k, name, shape, dt in ((0, "cnt", (s,), np.int32), (1, "off", (s,), np.int32), (2, "pos", (s, 3), np.int32),
                                   (3, "nid", (s,), np.int32), (4, "pnt", (length, 3), np.float32)):
            a = np.zeros(shape, dt)
            if a.size:
                L.orc_subcell_copy(self.h, k, _p(a))
            out[name] = a
        return out

    def gather_density(self, radius, scs=4, precision=32, accumulate=False):
        return lib().orc_gather_density(self.h, scs, float(radius), precision, int(accumulate))

    def gather_level_set(self, radius, scs=4, precision=32, accumulate=False):
        """GatherLevelSet: min(3, min over the subcell's entries of |p - v| - r)."""
        return lib().orc_gather_level_set(self.h, scs, float(radius), precision, int(accumulate))

    def compute(self, effect, iterations, params, update_apron, skip_aprons=True, boundval=0.0, precision=32):
        """Compute(effect, ...): fill 110, smooth 113, noise 114, grow 115."""
        p = [float(x) for x in params]
        return lib().orc_compute(self.h, int(effect), int(iterations), p[0], p[1], p[2], int(update_apron),
                                 int(skip_aprons), float(boundval), precision)

    def set_atlas(self, a, precision=32):
        a = np.ascontiguousarray(a, np.float64 if precision == 64 else np.float32)
        assert a.size == np.prod(self.atlas_info()["res"])
        lib().orc_atlas_set(self.h, precision, _p(a))

    def insert_points_subcell_fp16(self, pts, radius, scs=4, trans=(0, 0, 0)):
        """InsertPointsSubcell_FP16: the subcell lists under the FP16 binning
        rule, entries quantised to u16 against the point bounds."""
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        n = ctypes.c_int(0)
        rc = lib().orc_insert_points_subcell_fp16(self.h, scs, _p(pts), len(pts), float(radius), _p(f3(trans)),
                                                  ctypes.byref(n))
        return rc, n.value

    def subcells_fp16(self, length):
        a = np.zeros((length, 3), np.uint16)
        if a.size:
            lib().orc_subcell16_copy(self.h, _p(a))
        return a

    def gather_level_set_fp16(self, radius, scs=4, precision=32):
        return lib().orc_gather_level_set_fp16(self.h, scs, float(radius), precision)

    def update_apron(self, boundval=0.0, precision=32):
        return lib().orc_update_apron(self.h, float(boundval), precision)

    # render ---------------------------------------------------------------
    def render(self, scn, tf, shading, precision=64, capped=False):
        """Float RGBA image; with capped=True also a (h, w) bool mask of the
        rays that stopped at the reference's MAX_ITER walk cap."""
        scn = np.ascontiguousarray(scn, dtype=np.float32)
        w, h = int(scn[:2].view(np.int32)[0]), int(scn[:2].view(np.int32)[1])
        out = np.zeros((h, w, 4), np.float32)
        cap = np.zeros((h, w), np.uint8)
        tf = np.ascontiguousarray(tf, dtype=np.float32)
        lib().orc_render(self.h, _p(scn), _p(tf), int(shading), int(precision), _p(out), _p(cap))
        return (out, cap.astype(bool)) if capped else out

    def stage_ms(self, i):
        return lib().orc_stage_ms(self.h, i)


def linear_tf(t0, t1, a, b):
    lut = np.zeros((TF_SIZE, 4), np.float32)
    lib().orc_linear_tf(float(t0), float(t1), _p(np.asarray(a, np.float32)), _p(np.asarray(b, np.float32)), _p(lut))
    return lut


def scene_block(scene):
    """Scene dict (see workloads.ea_scene) -> the 68-float scene block."""
    cam = scene["cam"]
    c = np.asarray(list(cam["angs"]) + list(cam["to"]) + [cam["dist"], cam["dolly"], cam["fov"], cam["near"], cam["far"]],
                   np.float32)
    s = np.zeros(SCN_FLOATS, np.float32)
    lib().orc_camera_scn(_p(c), int(scene["width"]), int(scene["height"]), _p(s))
    s[14:17] = scene["light"]
    s[17:21] = scene["backclr"]
    s[21:24] = scene["extinct"]
    s[24:27] = scene["steps"]
    s[27:30] = scene["cutoff"]
    s[30:33] = scene["vrange"]
    s[33:36] = scene["shadow"]
    s[36:52] = np.eye(4, dtype=np.float32).reshape(-1)
    s[52:68] = np.eye(4, dtype=np.float32).reshape(-1)
    return s


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()
