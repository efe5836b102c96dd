"""Python binding of libgvdb_amd (ctypes over the C ABI in include/gvdb_amd.h).

`VolumeGVDB` mirrors the reference's nvdb::VolumeGVDB methods on the point
path (source/gvdb_library/src/gvdb_volume_gvdb.h:319-901 of NVIDIA GVDB
Voxels 1.1.1) with the same names and argument meaning.  Methods raise
GVDBError on a negative status; the library has no CPU fallback, so a missing
shared library or device is an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libgvdb_amd.so")

OK = 0
ERR_INVALID, ERR_HIP, ERR_STATE, ERR_DEPTH, ERR_NOBRICKS, ERR_RANGE, ERR_UNSUPPORTED, ERR_OOM = -1, -2, -3, -4, -5, -6, -7, -8
T_UCHAR4, T_FLOAT, T_HALF = 2, 3, 100
CHAN_UNDEF = 255
SHADE_TRILINEAR, SHADE_VOLUME = 4, 7
FUNC_FILL_F, FUNC_SMOOTH, FUNC_NOISE, FUNC_GROW = 110, 113, 114, 115
SCN_FLOATS = 68
# launch-shape knobs of the autotuner (gvdb_set_tuning / gvdb_get_tuning)
KNOB_RENDER_UNROLL_EA, KNOB_RENDER_UNROLL_ISO, KNOB_RENDER_TILE_EA, KNOB_RENDER_TILE_ISO = 0, 1, 2, 3
KNOB_GATHER_THREADS, KNOB_APRON_RUN, KNOB_XCD_ORDER, KNOB_APRON_QUEUE, KNOB_MARK_PATH = 4, 5, 6, 7, 8
KNOB_APRON_RUN_ONE, KNOB_COUNT = 9, 10
# the allocation-free UpdateApron: size of its fixed device array and the bricks it holds
NOALLOC_SCRATCH_BYTES = 32 << 20
NOALLOC_MAX_BRICKS = NOALLOC_SCRATCH_BYTES // (27 * 12)

NODE_DTYPE = np.dtype([
    ("mLev", "u1"), ("mFlags", "u1"), ("mPriority", "u1"), ("pad", "u1"),
    ("mPos", "<i4", 3), ("mValue", "<i4", 3), ("mVRange", "<f4", 3),
    ("mParent", "<u8"), ("mChildList", "<u8"), ("mMask", "<u8")])
ATLASNODE_DTYPE = np.dtype([("mPos", "<i4", 3), ("mLeafNode", "<i4")])

# (name, restype, argtypes) of every C entry point
_vp, _i, _f, _u64p = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_uint64)
_SIGS = [
    ("gvdb_create", _i, [ctypes.POINTER(_vp), _i]), ("gvdb_destroy", _i, [_vp]),
    ("gvdb_last_error", ctypes.c_char_p, [_vp]),
    ("gvdb_configure", _i, [_vp, _i, _vp, _vp]), ("gvdb_set_channel_default", _i, [_vp, _i, _i, _i]),
    ("gvdb_add_channel", _i, [_vp, _i, _i, _i, _i, _i, _i, _i, _i]), ("gvdb_set_epsilon", _i, [_vp, _f, _i]),
    ("gvdb_set_points_gpu", _i, [_vp, _vp, _i, _i, _i]), ("gvdb_upload_points", _i, [_vp, _vp, _i]),
    ("gvdb_set_colors_gpu", _i, [_vp, _vp, _i, _i, _i]), ("gvdb_upload_colors", _i, [_vp, _vp, _i]),
    ("gvdb_set_color_channel", _i, [_vp, _i]),
    ("gvdb_request_full_rebuild", _i, [_vp, _i]), ("gvdb_rebuild_topology", _i, [_vp, _i, _f, _vp]),
    ("gvdb_rebuild_topology_cpu", _i, [_vp, _i, _vp, _vp]), ("gvdb_activate_space", _i, [_vp, _vp]),
    ("gvdb_accumulate_topology", _i, [_vp, _i, _f, _vp, _i]),
    ("gvdb_save_vbx", _i, [_vp, ctypes.c_char_p]), ("gvdb_load_vbx", _i, [_vp, ctypes.c_char_p, _i, _i]),
    ("gvdb_get_transform", _i, [_vp, _vp]), ("gvdb_set_transform", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("gvdb_compute", _i, [_vp, _i, _i, _i, _vp, _i, _i, _f]),
    ("gvdb_compute_kernel", _i, [_vp, _vp, _i, _i, _i]), ("gvdb_render_kernel", _i, [_vp, _vp, _i, _i]),
    ("gvdb_module_load", _i, [ctypes.c_char_p, _vp]), ("gvdb_module_function", _i, [_vp, ctypes.c_char_p, _vp]),
    ("gvdb_module_unload", _i, [_vp]),
    ("gvdb_clear", _i, [_vp]), ("gvdb_finish_topology", _i, [_vp, _i, _i]), ("gvdb_update_atlas", _i, [_vp]),
    ("gvdb_clear_channel", _i, [_vp, _i]),
    ("gvdb_insert_points_subcell", _i, [_vp, _i, _i, _f, _vp, ctypes.POINTER(_i)]),
    ("gvdb_gather_density", _i, [_vp, _i, _i, _f, _vp, _i, _i, _i, _i]), ("gvdb_update_apron", _i, [_vp, _i, _f]),
    ("gvdb_update_apron_noalloc", _i, [_vp, _i, _f]),
    ("gvdb_gather_level_set", _i, [_vp, _i, _i, _f, _vp, _i, _i, _i, _i]),
    ("gvdb_insert_points_subcell_fp16", _i, [_vp, _i, _i, _f, _vp, _vp]),
    ("gvdb_gather_level_set_fp16", _i, [_vp, _i, _i, _f, _vp, _i, _i, _i]),
    ("gvdb_set_camera_orbit", _i, [_vp, _vp, _vp, _f, _f, _f, _f, _f]), ("gvdb_set_light", _i, [_vp, _vp]),
    ("gvdb_set_background", _i, [_vp, _vp]), ("gvdb_set_extinct", _i, [_vp, _f, _f, _f]),
    ("gvdb_set_steps", _i, [_vp, _f, _f, _f]), ("gvdb_set_cutoff", _i, [_vp, _f, _f, _f]),
    ("gvdb_set_volume_range", _i, [_vp, _f, _f, _f]), ("gvdb_set_shadow_params", _i, [_vp, _f, _f, _f]),
    ("gvdb_linear_transfer_func", _i, [_vp, _f, _f, _vp, _vp]), ("gvdb_commit_transfer_func", _i, [_vp]),
    ("gvdb_add_render_buf", _i, [_vp, _i, _i, _i, _i]), ("gvdb_render", _i, [_vp, _i, _i, _i]),
    ("gvdb_read_render_buf", _i, [_vp, _i, _vp]), ("gvdb_get_scene_block", _i, [_vp, _i, _i, _vp]),
    ("gvdb_tune_render", _i, [_vp, _i, _i, _i, _vp, _vp]), ("gvdb_set_render_unroll", _i, [_i, _i]),
    ("gvdb_tune_render_shape", _i, [_vp, _i, _i, _i, _vp, _vp]),
    ("gvdb_tune_gather", _i, [_vp, _i, _i, _f, _vp, _i, _i, _vp, _vp]),
    ("gvdb_set_tuning", _i, [_i, _i]), ("gvdb_get_tuning", _i, [_i, _vp, _vp, _vp]),
    ("gvdb_capture_begin", _i, [_vp, ctypes.POINTER(_vp)]), ("gvdb_capture_end", _i, [_vp, _vp]),
    ("gvdb_batch_capture_begin", _i, [_vp, ctypes.POINTER(_vp)]), ("gvdb_batch_capture_end", _i, [_vp, _vp]),
    ("gvdb_graph_launch", _i, [_vp]), ("gvdb_graph_synchronize", _i, [_vp]), ("gvdb_graph_destroy", _i, [_vp]),
    ("gvdb_graph_last_error", ctypes.c_char_p, [_vp]),
    ("gvdb_batch_tune", _i, [_vp, _i, _vp, _i, _i, _vp, _vp, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("gvdb_get_pool", _i, [_vp, _i, _i, _vp, _u64p, _u64p]), ("gvdb_get_info", _i, [_vp, _vp, _vp]),
    ("gvdb_get_atlas_info", _i, [_vp, _i, _vp]), ("gvdb_read_atlas", _i, [_vp, _i, _vp]),
    ("gvdb_read_atlas_map", _i, [_vp, _vp]), ("gvdb_read_subcells", _i, [_vp, _i, _vp, _u64p]),
    ("gvdb_get_stream", _i, [_vp, ctypes.POINTER(_vp)]), ("gvdb_synchronize", _i, [_vp]),
    ("gvdb_stage_ms", _i, [_vp, _vp]), ("gvdb_kernel_ms", _i, [_vp, _vp]),
    ("gvdb_batch_create", _i, [ctypes.POINTER(_vp), _i, _i]), ("gvdb_batch_destroy", _i, [_vp]),
    ("gvdb_batch_last_error", ctypes.c_char_p, [_vp]),
    ("gvdb_batch_rebuild_topology", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    ("gvdb_batch_finish_topology", _i, [_vp, _i, _vp, _i, _i, _vp]),
    ("gvdb_batch_update_atlas", _i, [_vp, _i, _vp, _vp]), ("gvdb_batch_clear_channel", _i, [_vp, _i, _vp, _i, _vp]),
    ("gvdb_batch_insert_points_subcell", _i, [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    ("gvdb_batch_gather_density", _i, [_vp, _i, _vp, _i, _vp, _vp, _i, _i, _vp]),
    ("gvdb_batch_gather_level_set", _i, [_vp, _i, _vp, _i, _vp, _vp, _i, _i, _vp]),
    ("gvdb_batch_update_apron", _i, [_vp, _i, _vp, _i, _f, _vp]),
    ("gvdb_batch_render", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    ("gvdb_batch_synchronize", _i, [_vp]), ("gvdb_batch_elapsed_ms", _i, [_vp, _vp]),
    ("gvdb_release_stream", _i, [_vp]),
    ("gvdb_server_create", _i, [ctypes.POINTER(_vp), _i, ctypes.c_ulonglong, _i]),
    ("gvdb_server_destroy", _i, [_vp]), ("gvdb_server_last_error", ctypes.c_char_p, [_vp]),
    ("gvdb_server_add_grid", _i, [_vp, _vp, _vp, _i, _f]), ("gvdb_server_reset", _i, [_vp]),
    ("gvdb_server_submit_render", _i, [_vp, _i, ctypes.c_double, _i, _i, _i, _vp, _f, _vp]),
    ("gvdb_server_submit_insert", _i, [_vp, _i, ctypes.c_double, _vp, _i]),
    ("gvdb_server_clear_requests", _i, [_vp]), ("gvdb_server_run", _i, [_vp, _i, _vp, _vp, _vp]),
    ("gvdb_server_device_bytes", ctypes.c_ulonglong, [_vp]),
]

_lib = None


class GVDBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def load():
    """Loads libgvdb_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libgvdb_amd.so not built: run gvdb-voxels_amd/build.py (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_names():
    return [s[0] for s in _SIGS]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


def get_tuning(knob):
    """(current value, values it can take) of a launch-shape knob."""
    v, n = ctypes.c_int(0), ctypes.c_int(0)
    ch = np.zeros(8, np.int32)
    rc = load().gvdb_get_tuning(int(knob), ctypes.byref(v), _p(ch), ctypes.byref(n))
    if rc != OK:
        raise GVDBError(rc, "gvdb_get_tuning(%d)" % knob)
    return v.value, [int(x) for x in ch[:n.value]]


def set_tuning(knob, value):
    rc = load().gvdb_set_tuning(int(knob), int(value))
    if rc != OK:
        raise GVDBError(rc, "gvdb_set_tuning(%d, %d)" % (knob, value))


def module_load(path):
    """hipModuleLoad of a code object; returns the module handle."""
    m = ctypes.c_void_p()
    rc = load().gvdb_module_load(os.fsencode(path), ctypes.byref(m))
    if rc:
        raise GVDBError(rc, "module_load: " + str(path))
    return m.value


def module_function(module, name):
    f = ctypes.c_void_p()
    rc = load().gvdb_module_function(ctypes.c_void_p(module), name.encode(), ctypes.byref(f))
    if rc:
        raise GVDBError(rc, "module_function: " + name)
    return f.value


def module_unload(module):
    load().gvdb_module_unload(ctypes.c_void_p(module))


def user_kernels_path():
    """The sample user kernels built with the library."""
    return os.path.join(os.path.dirname(LIB_PATH), "user_kernels.hsaco")


class VolumeGVDB:
    """One grid on one GPU (VolumeGVDB + Initialize)."""

    def __init__(self, device=0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.gvdb_create(ctypes.byref(h), int(device))
        if rc != OK:
            raise GVDBError(rc, "gvdb_create failed (no HIP device %d?)" % device)
        self.h = h
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            load().gvdb_destroy(self.h)
            self.h = None

    def close(self):
        self.__del__()

    def _chk(self, rc, allow=()):
        if rc != OK and rc not in allow:
            raise GVDBError(rc, load().gvdb_last_error(self.h).decode())
        return rc

    # setup -----------------------------------------------------------------
    def Configure(self, q4, q3, q2, q1, q0):
        dims = np.asarray([q0, q1, q2, q3, q4], np.int32)
        return self._chk(load().gvdb_configure(self.h, 5, _p(dims), None))

    def ConfigureLevels(self, log2dims, initcnt=None):
        dims = np.asarray(log2dims, np.int32)
        cnt = None if initcnt is None else _p(np.asarray(initcnt, np.int32))
        return self._chk(load().gvdb_configure(self.h, len(dims), _p(dims), cnt))

    def SetChannelDefault(self, x, y, z):
        return self._chk(load().gvdb_set_channel_default(self.h, x, y, z))

    def AddChannel(self, chan, dtype=T_FLOAT, apron=1, filter=0, border=0, axiscnt=(0, 0, 0)):
        return self._chk(load().gvdb_add_channel(self.h, chan, dtype, apron, filter, border, *axiscnt))

    def SetEpsilon(self, eps, maxiter=256):
        return self._chk(load().gvdb_set_epsilon(self.h, float(eps), int(maxiter)))

    # points ----------------------------------------------------------------
    def SetPoints(self, pts, colors=None):
        """Uploads host float3 positions (AllocData + CommitData + SetPoints),
        and with `colors` a uint32 (uchar4 r, g, b, a) per point; without,
        the colours are cleared."""
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        self._chk(load().gvdb_upload_points(self.h, _p(pts), len(pts)))
        if colors is None:
            return self._chk(load().gvdb_set_colors_gpu(self.h, None, 0, 0, 4))
        c = np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
        return self._chk(load().gvdb_upload_colors(self.h, _p(c), len(c)))

    def SetColorChannel(self, chan):
        """SetColorChannel(chan): the march reads that uchar4 channel's colours (255 or -1: none)."""
        return self._chk(load().gvdb_set_color_channel(self.h, int(chan)))

    def SetDataGPU(self, dev_ptr, n, offset=0, stride=12):
        return self._chk(load().gvdb_set_points_gpu(self.h, ctypes.c_void_p(dev_ptr), n, offset, stride))

    # topology --------------------------------------------------------------
    def RequestFullRebuild(self, on=True):
        return self._chk(load().gvdb_request_full_rebuild(self.h, int(on)))

    def RebuildTopology(self, n, radius, orig=(0, 0, 0)):
        return self._chk(load().gvdb_rebuild_topology(self.h, int(n), float(radius), _p(_f3(orig))))

    def AccumulateTopology(self, n, radius, orig=(0, 0, 0), depth=1):
        return self._chk(load().gvdb_accumulate_topology(self.h, int(n), float(radius), _p(_f3(orig)), int(depth)))

    def RebuildTopologyCPU(self, pts, orig=(0, 0, 0)):
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        return self._chk(load().gvdb_rebuild_topology_cpu(self.h, len(pts), _p(_f3(orig)), _p(pts)))

    def ActivateSpace(self, pos):
        return self._chk(load().gvdb_activate_space(self.h, _p(_f3(pos))))

    def Clear(self):
        return self._chk(load().gvdb_clear(self.h))

    def FinishTopology(self, commit=True, bounds=True):
        return self._chk(load().gvdb_finish_topology(self.h, int(commit), int(bounds)))

    def UpdateAtlas(self):
        return self._chk(load().gvdb_update_atlas(self.h))

    def ClearChannel(self, chan):
        return self._chk(load().gvdb_clear_channel(self.h, chan))

    # points -> voxels ------------------------------------------------------
    def InsertPointsSubcell(self, subcell_size, n, radius, trans=(0, 0, 0)):
        ln = ctypes.c_int(0)
        self._chk(load().gvdb_insert_points_subcell(self.h, subcell_size, int(n), float(radius), _p(_f3(trans)),
                                                    ctypes.byref(ln)))
        return ln.value

    def GatherDensity(self, subcell_size, n, radius, trans, sc_len, chan_density=0, chan_clr=0, accumulate=False):
        return self._chk(load().gvdb_gather_density(self.h, subcell_size, int(n), float(radius), _p(_f3(trans)),
                                                    int(sc_len), chan_density, chan_clr, int(accumulate)))

    def GatherLevelSet(self, subcell_size, n, radius, trans, sc_len, chan_levelset=0, chan_clr=0, accumulate=False):
        return self._chk(load().gvdb_gather_level_set(self.h, subcell_size, int(n), float(radius), _p(_f3(trans)),
                                                      int(sc_len), chan_levelset, chan_clr, int(accumulate)))

    def Compute(self, effect, chan, iterations, params, update_apron, skip_aprons=True, boundval=0.0):
        return self._chk(load().gvdb_compute(self.h, int(effect), int(chan), int(iterations), _p(_f3(params)),
                                             int(update_apron), int(skip_aprons), float(boundval)))

    def ComputeKernel(self, function, chan, update_apron, skip_aprons=True):
        """Launches a user kernel (a hipFunction_t from module_function) over the atlas."""
        return self._chk(load().gvdb_compute_kernel(self.h, ctypes.c_void_p(function), int(chan), int(update_apron),
                                                    int(skip_aprons)))

    def RenderKernel(self, function, chan=0, rbuf=0):
        return self._chk(load().gvdb_render_kernel(self.h, ctypes.c_void_p(function), int(chan), int(rbuf)))

    def SaveVBX(self, path):
        return self._chk(load().gvdb_save_vbx(self.h, os.fsencode(path)))

    def LoadVBX(self, path, force_major=0, force_minor=0):
        return self._chk(load().gvdb_load_vbx(self.h, os.fsencode(path), int(force_major), int(force_minor)))

    def SetTransform(self, pretrans, scale, angs, trans):
        """SetTransform: the grid placed in application space as T R S PT."""
        return self._chk(load().gvdb_set_transform(self.h, _p(_f3(pretrans)), _p(_f3(scale)), _p(_f3(angs)),
                                                   _p(_f3(trans))))

    def transform(self):
        """(pretrans, angs, scale, trans) of the grid, as a VBX file stores them."""
        out = np.zeros(12, np.float32)
        self._chk(load().gvdb_get_transform(self.h, _p(out)))
        return out.reshape(4, 3)

    def InsertPointsSubcell_FP16(self, subcell_size, n, radius, trans=(0, 0, 0)):
        ln = ctypes.c_int(0)
        self._chk(load().gvdb_insert_points_subcell_fp16(self.h, subcell_size, int(n), float(radius), _p(_f3(trans)),
                                                         ctypes.byref(ln)))
        return ln.value

    def GatherLevelSet_FP16(self, subcell_size, n, radius, trans, sc_len, chan_levelset=0, chan_clr=0):
        return self._chk(load().gvdb_gather_level_set_fp16(self.h, subcell_size, int(n), float(radius), _p(_f3(trans)),
                                                           int(sc_len), chan_levelset, chan_clr))

    def UpdateApron(self, chan=0, boundval=0.0):
        return self._chk(load().gvdb_update_apron(self.h, chan, float(boundval)))

    def UpdateApronNoAlloc(self, chan=0, boundval=0.0):
        """UpdateApron with the neighbour table in the library's fixed device
        array: allocates nothing; at most NOALLOC_MAX_BRICKS bricks."""
        return self._chk(load().gvdb_update_apron_noalloc(self.h, chan, float(boundval)))

    # scene / render --------------------------------------------------------
    def apply_scene(self, sc):
        L = load()
        c = sc["cam"]
        self._chk(L.gvdb_set_camera_orbit(self.h, _p(_f3(c["angs"])), _p(_f3(c["to"])), float(c["dist"]),
                                          float(c["dolly"]), float(c["fov"]), float(c["near"]), float(c["far"])))
        self._chk(L.gvdb_set_light(self.h, _p(_f3(sc["light"]))))
        self._chk(L.gvdb_set_background(self.h, _p(np.asarray(sc["backclr"], np.float32))))
        self._chk(L.gvdb_set_extinct(self.h, *map(float, sc["extinct"])))
        self._chk(L.gvdb_set_steps(self.h, *map(float, sc["steps"])))
        self._chk(L.gvdb_set_cutoff(self.h, *map(float, sc["cutoff"])))
        self._chk(L.gvdb_set_volume_range(self.h, *map(float, sc["vrange"])))
        self._chk(L.gvdb_set_shadow_params(self.h, *map(float, sc["shadow"])))
        t0, t1, a, b = sc["tf"]
        self._chk(L.gvdb_linear_transfer_func(self.h, float(t0), float(t1), _p(np.asarray(a, np.float32)),
                                              _p(np.asarray(b, np.float32))))
        return self._chk(L.gvdb_commit_transfer_func(self.h))

    def AddRenderBuf(self, buf, w, h, bytes_per_pixel=4):
        return self._chk(load().gvdb_add_render_buf(self.h, buf, w, h, bytes_per_pixel))

    def Render(self, shading, chan=0, buf=0):
        return self._chk(load().gvdb_render(self.h, shading, chan, buf))

    @staticmethod
    def SetRenderUnroll(shading, unroll):
        """Samples fetched ahead per brick step (1, 2, 4) for a shading mode, process-wide."""
        rc = load().gvdb_set_render_unroll(int(shading), int(unroll))
        if rc != 0:
            raise GVDBError(rc, "gvdb_set_render_unroll(%d, %d)" % (shading, unroll))

    def TuneRender(self, shading, chan=0, buf=0):
        """Autotunes the samples fetched ahead per brick step for this shading
        mode on the current scene; returns (best, {unroll: ms})."""
        best = ctypes.c_int(0)
        ms = np.zeros(3, np.float32)
        self._chk(load().gvdb_tune_render(self.h, int(shading), int(chan), int(buf), ctypes.byref(best), _p(ms)))
        return best.value, {1: float(ms[0]), 2: float(ms[1]), 4: float(ms[2])}

    def CaptureBegin(self):
        """Starts recording the per-frame chain of this grid into a graph."""
        g = ctypes.c_void_p()
        self._chk(load().gvdb_capture_begin(self.h, ctypes.byref(g)))
        self._graph = Graph(g)

    def CaptureEnd(self):
        """Ends the recording; returns the Graph."""
        g, self._graph = self._graph, None
        rc = load().gvdb_capture_end(self.h, g.h)
        if rc != OK:
            msg = load().gvdb_graph_last_error(g.h).decode()
            g.close()
            raise GVDBError(rc, "capture: " + msg)
        return g

    def TuneRenderShape(self, shading, chan=0, buf=0):
        """Autotunes the ray-packet shape and the samples fetched ahead for
        this shading mode on the current scene; returns ((unroll, tile
        width), {(tile width, unroll): ms})."""
        best = np.zeros(2, np.int32)
        ms = np.zeros(12, np.float32)
        self._chk(load().gvdb_tune_render_shape(self.h, int(shading), int(chan), int(buf), _p(best), _p(ms)))
        return (int(best[0]), int(best[1])), {(tw, u): float(ms[i * 3 + j]) for i, tw in enumerate((8, 16, 32, 64))
                                              for j, u in enumerate((1, 2, 4))}

    def TuneGather(self, subcell_size, n, radius, trans, sc_len, chan=0):
        """Autotunes the gather's threads per brick (GatherDensity without
        accumulate, run on the current buckets); returns (best, {threads: ms})."""
        best = ctypes.c_int(0)
        ms = np.zeros(2, np.float32)
        self._chk(load().gvdb_tune_gather(self.h, int(subcell_size), int(n), float(radius), _p(_f3(trans)),
                                          int(sc_len), int(chan), ctypes.byref(best), _p(ms)))
        return best.value, {128: float(ms[0]), 256: float(ms[1])}

    def ReadRenderBuf(self, buf, w, h, bytes_per_pixel=4):
        out = np.zeros((h, w, 4), np.float32 if bytes_per_pixel == 16 else np.uint8)
        self._chk(load().gvdb_read_render_buf(self.h, buf, _p(out)))
        return out

    def scene_block(self, w, h):
        s = np.zeros(SCN_FLOATS, np.float32)
        self._chk(load().gvdb_get_scene_block(self.h, w, h, _p(s)))
        return s

    # access ----------------------------------------------------------------
    def pool(self, grp, lev):
        L = load()
        cnt, stride = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._chk(L.gvdb_get_pool(self.h, grp, lev, None, ctypes.byref(cnt), ctypes.byref(stride)))
        buf = np.zeros(int(cnt.value * stride.value), np.uint8)
        if cnt.value:
            self._chk(L.gvdb_get_pool(self.h, grp, lev, _p(buf), ctypes.byref(cnt), ctypes.byref(stride)))
        if grp == 0:
            return buf.view(NODE_DTYPE)
        return buf.view(np.uint64).reshape(int(cnt.value), -1) if cnt.value else np.zeros((0, 0), np.uint64)

    def info(self):
        i2 = np.zeros(2, np.int32)
        f = np.zeros(12, np.float32)
        self._chk(load().gvdb_get_info(self.h, _p(i2), _p(f)))
        return dict(top_lev=int(i2[0]), depth=int(i2[1]), obj_min=f[0:3], obj_max=f[3:6], pos_min=f[6:9],
                    pos_max=f[9:12])

    def atlas_info(self, chan=0):
        o = np.zeros(7, np.int32)
        self._chk(load().gvdb_get_atlas_info(self.h, chan, _p(o)))
        return dict(cnt=tuple(int(x) for x in o[0:3]), res=tuple(int(x) for x in o[3:6]), used=int(o[6]))

    def atlas(self, chan=0, dtype=np.float32):
        rx, ry, rz = self.atlas_info(chan)["res"]
        a = np.zeros((rz, ry, rx), dtype)
        self._chk(load().gvdb_read_atlas(self.h, chan, _p(a)))
        return a

    def atlas_map(self):
        n = int(np.prod(self.atlas_info(0)["cnt"]))
        a = np.zeros(n, ATLASNODE_DTYPE)
        self._chk(load().gvdb_read_atlas_map(self.h, _p(a)))
        return a

    def subcells(self):
        L = load()
        out = {}
        for k, name, cols, dt in ((0, "cnt", 1, np.uint32), (1, "off", 1, np.uint32), (2, "pos", 3, np.int32),
                                  (3, "nid", 1, np.int32), (4, "pnt", 3, np.float32), (5, "mapping", 1, np.uint32)):
            c = ctypes.c_uint64(0)
            self._chk(L.gvdb_read_subcells(self.h, k, None, ctypes.byref(c)))
            a = np.zeros((int(c.value), cols) if cols > 1 else int(c.value), dt)
            if c.value:
                self._chk(L.gvdb_read_subcells(self.h, k, _p(a), ctypes.byref(c)))
            out[name] = a
        return out

    def subcells_fp16(self):
        """u16 entries of InsertPointsSubcell_FP16, one row per entry."""
        L = load()
        c = ctypes.c_uint64(0)
        self._chk(L.gvdb_read_subcells(self.h, 6, None, ctypes.byref(c)))
        a = np.zeros((int(c.value), 3), np.uint16)
        if c.value:
            self._chk(L.gvdb_read_subcells(self.h, 6, _p(a), ctypes.byref(c)))
        return a

    def synchronize(self):
        return self._chk(load().gvdb_synchronize(self.h))

    def stage_ms(self):
        o = np.zeros(8, np.float32)
        self._chk(load().gvdb_stage_ms(self.h, _p(o)))
        return dict(zip(("topology", "atlas", "insert", "gather", "apron", "render", "clear"), map(float, o[:7])))

    def kernel_ms(self):
        o = np.zeros(4, np.float32)
        self._chk(load().gvdb_kernel_ms(self.h, _p(o)))
        return dict(gather=float(o[0]), apron=float(o[1]), render=float(o[2]))


class Graph:
    """A recorded per-frame chain (gvdb_capture_* / gvdb_batch_capture_*):
    launch() replays it, synchronize() waits and brings the insert's entry
    count back."""

    def __init__(self, handle):
        self.h = handle

    def _chk(self, rc):
        if rc != OK:
            raise GVDBError(rc, "graph: " + load().gvdb_graph_last_error(self.h).decode())
        return rc

    def launch(self):
        return self._chk(load().gvdb_graph_launch(self.h))

    def synchronize(self):
        return self._chk(load().gvdb_graph_synchronize(self.h))

    def close(self):
        if self.h:
            load().gvdb_graph_destroy(self.h)
            self.h = None


class GridBatch:
    """Many grids per launch: each method runs one stage of the point path
    over a list of VolumeGVDB objects (any sizes) with one kernel launch per
    kernel of the stage.  Per-grid arguments are sequences, or one value for
    all."""

    def __init__(self, device=0, max_grids=256):
        L = load()
        h = ctypes.c_void_p()
        rc = L.gvdb_batch_create(ctypes.byref(h), int(device), int(max_grids))
        if rc != OK:
            raise GVDBError(rc, "gvdb_batch_create failed (no HIP device %d?)" % device)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            load().gvdb_batch_destroy(self.h)
            self.h = None

    def close(self):
        self.__del__()

    def _chk(self, rc, vols=None, failed=None):
        if rc != OK:
            who = int(failed.value) if failed is not None else -1
            raise GVDBError(rc, "batch (grid %d): %s" % (who, load().gvdb_batch_last_error(self.h).decode()))
        return rc

    @staticmethod
    def _hs(vols):
        return (ctypes.c_void_p * len(vols))(*[v.h.value for v in vols])

    @staticmethod
    def _arr(x, G, dt):
        a = np.asarray(x, dt)
        if a.ndim == 0:
            a = np.full(G, a, dt)
        return np.ascontiguousarray(a.reshape(G))

    @staticmethod
    def _arr3(x, G):
        if x is None:
            return np.zeros((G, 3), np.float32)
        a = np.asarray(x, np.float32)
        if a.shape == (3,):
            a = np.tile(a, (G, 1))
        return np.ascontiguousarray(a.reshape(G, 3))

    def RebuildTopology(self, vols, n, radius, orig=None):
        G = len(vols)
        n, r, o = self._arr(n, G, np.int32), self._arr(radius, G, np.float32), self._arr3(orig, G)
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_rebuild_topology(self.h, G, self._hs(vols), _p(n), _p(r), _p(o),
                                                            ctypes.byref(f)), vols, f)

    def FinishTopology(self, vols, commit=False, bounds=True):
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_finish_topology(self.h, len(vols), self._hs(vols), int(commit), int(bounds),
                                                           ctypes.byref(f)), vols, f)

    def UpdateAtlas(self, vols):
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_update_atlas(self.h, len(vols), self._hs(vols), ctypes.byref(f)), vols, f)

    def ClearChannel(self, vols, chan=0):
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_clear_channel(self.h, len(vols), self._hs(vols), chan, ctypes.byref(f)),
                         vols, f)

    def InsertPointsSubcell(self, vols, subcell_size, n, radius, trans=None):
        G = len(vols)
        n, r, t = self._arr(n, G, np.int32), self._arr(radius, G, np.float32), self._arr3(trans, G)
        out = np.zeros(G, np.int32)
        f = ctypes.c_int(-1)
        self._chk(load().gvdb_batch_insert_points_subcell(self.h, G, self._hs(vols), subcell_size, _p(n), _p(r), _p(t),
                                                          _p(out), ctypes.byref(f)), vols, f)
        return [int(x) for x in out]

    def GatherDensity(self, vols, subcell_size, sc_len, radius, chan_density=0, accumulate=False):
        G = len(vols)
        ln, r = self._arr(sc_len, G, np.int32), self._arr(radius, G, np.float32)
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_gather_density(self.h, G, self._hs(vols), subcell_size, _p(ln), _p(r),
                                                          chan_density, int(accumulate), ctypes.byref(f)), vols, f)

    def GatherLevelSet(self, vols, subcell_size, sc_len, radius, chan_levelset=0, accumulate=False):
        G = len(vols)
        ln, r = self._arr(sc_len, G, np.int32), self._arr(radius, G, np.float32)
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_gather_level_set(self.h, G, self._hs(vols), subcell_size, _p(ln), _p(r),
                                                            chan_levelset, int(accumulate), ctypes.byref(f)), vols, f)

    def UpdateApron(self, vols, chan=0, boundval=0.0):
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_update_apron(self.h, len(vols), self._hs(vols), chan, float(boundval),
                                                        ctypes.byref(f)), vols, f)

    def Render(self, vols, shading, chan=0, buf=0):
        G = len(vols)
        sh, c, b = self._arr(shading, G, np.int32), self._arr(chan, G, np.int32), self._arr(buf, G, np.int32)
        f = ctypes.c_int(-1)
        return self._chk(load().gvdb_batch_render(self.h, G, self._hs(vols), _p(sh), _p(c), _p(b), ctypes.byref(f)),
                         vols, f)

    def Tune(self, vols, stage, subcell_size=4, sc_len=None, radius=None, chan=0, boundval=0.0, shading=SHADE_VOLUME,
             buf=0):
        """Batched autotuner: stage "gather" (threads per brick), "apron"
        (bricks per workgroup) or "render" (ray-packet shape x samples
        ahead), timed over the whole batch; returns (best, {value: ms})."""
        G = len(vols)
        k = {"gather": 0, "apron": 1, "render": 2}[stage]
        ln = self._arr(sc_len if sc_len is not None else 0, G, np.int32)
        r = self._arr(radius if radius is not None else 0.0, G, np.float32)
        sh, c, b = self._arr(shading, G, np.int32), self._arr(chan, G, np.int32), self._arr(buf, G, np.int32)
        best = np.zeros(2, np.int32)
        ms = np.zeros(12, np.float32)
        f = ctypes.c_int(-1)
        self._chk(load().gvdb_batch_tune(self.h, G, self._hs(vols), k, subcell_size, _p(ln), _p(r), chan,
                                         float(boundval), _p(sh), _p(c), _p(b), _p(best), _p(ms), ctypes.byref(f)),
                  vols, f)
        if k == 0:
            return int(best[0]), {128: float(ms[0]), 256: float(ms[1])}
        if k == 1:
            d = {v: float(ms[i]) for i, v in enumerate((1, 2, 4, 8))}
            d["queue"] = float(ms[4])
            return ("queue" if best[1] else int(best[0])), d
        return (int(best[0]), int(best[1])), {(tw, u): float(ms[i * 3 + j]) for i, tw in enumerate((8, 16, 32, 64))
                                              for j, u in enumerate((1, 2, 4))}

    def CaptureBegin(self):
        """Starts recording the batched per-frame chain into a graph (every
        grid must have taken part in a batched call before)."""
        g = ctypes.c_void_p()
        self._chk(load().gvdb_batch_capture_begin(self.h, ctypes.byref(g)))
        self._graph = Graph(g)

    def CaptureEnd(self):
        g, self._graph = self._graph, None
        rc = load().gvdb_batch_capture_end(self.h, g.h)
        if rc != OK:
            msg = load().gvdb_graph_last_error(g.h).decode()
            g.close()
            raise GVDBError(rc, "batch capture: " + msg)
        return g

    def synchronize(self):
        return self._chk(load().gvdb_batch_synchronize(self.h))

    def elapsed_ms(self):
        ms = np.zeros(1, np.float32)
        self._chk(load().gvdb_batch_elapsed_ms(self.h, _p(ms)))
        return float(ms[0])


class Server:
    """Serving loop over resident grids (see include/gvdb_amd.h): add grids,
    submit timed render / insert requests, run them plainly (mode 0) or with
    continuous batching (mode 1)."""

    PLAIN, CONTINUOUS = 0, 1

    def __init__(self, device=0, hbm_budget=200 << 30, max_batch=64):
        L = load()
        h = ctypes.c_void_p()
        rc = L.gvdb_server_create(ctypes.byref(h), int(device), int(hbm_budget), int(max_batch))
        if rc != OK:
            raise GVDBError(rc, "gvdb_server_create failed")
        self.h = h
        self._keep = []
        self.requests = []

    def __del__(self):
        if getattr(self, "h", None):
            load().gvdb_server_destroy(self.h)
            self.h = None

    def close(self):
        self.__del__()

    def _chk(self, rc):
        if rc < 0:
            raise GVDBError(rc, load().gvdb_server_last_error(self.h).decode())
        return rc

    def add_grid(self, vol, pts, radius=2.0):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        self._keep.append(vol)
        return self._chk(load().gvdb_server_add_grid(self.h, vol.h, _p(pts), len(pts), float(radius)))

    def reset(self):
        return self._chk(load().gvdb_server_reset(self.h))

    def submit_render(self, grid, arrival, w, h, shading=SHADE_VOLUME, angs=(30.0, 20.0, 0.0), dist=0.0):
        out = np.zeros((h, w, 4), np.uint8)
        a = np.asarray(angs, np.float32)
        i = self._chk(load().gvdb_server_submit_render(self.h, int(grid), float(arrival), int(w), int(h), int(shading),
                                                       _p(a), float(dist), _p(out)))
        self._keep.append(out)
        self.requests.append(("render", grid, out))
        return i

    def submit_insert(self, grid, arrival, pts):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        self._keep.append(pts)
        i = self._chk(load().gvdb_server_submit_insert(self.h, int(grid), float(arrival), _p(pts), len(pts)))
        self.requests.append(("insert", grid, pts))
        return i

    def clear_requests(self):
        self.requests = []
        return self._chk(load().gvdb_server_clear_requests(self.h))

    def run(self, mode):
        n = len(self.requests)
        lat = np.zeros(max(n, 1), np.float64)
        st = np.zeros(max(n, 1), np.int32)
        wall = np.zeros(1, np.float64)
        self._chk(load().gvdb_server_run(self.h, int(mode), _p(lat), _p(st), _p(wall)))
        return lat[:n], st[:n], float(wall[0])

    def device_bytes(self):
        return int(load().gvdb_server_device_bytes(self.h))
