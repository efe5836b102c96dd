"""The C ABI library on a host without a GPU: it loads, exports every entry
point that include/gvdb_amd.h declares, agrees with the Python binding and the
header on codes and record layouts, and fails loudly (no CPU fallback) when no
HIP device is present.  The CPU tests make no compute call; the one GPU test
runs the C++ facade sample against the Python binding."""
import ctypes
import os
import re
import subprocess

import numpy as np

import helpers as H
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gvdb_amd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*|unsigned long long)\s+(gvdb_\w+)\s*\(", txt, re.M)))


def _defines():
    txt = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"#define\s+(GVDB_\w+)\s+\(?(-?\d+)\)?", txt)}


@pytest.fixture(scope="module")
def lib():
    import gvdb_amd
    return gvdb_amd.load()


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 40
    for n in ("gvdb_create", "gvdb_rebuild_topology", "gvdb_insert_points_subcell", "gvdb_gather_density",
              "gvdb_update_apron", "gvdb_render", "gvdb_get_pool"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, "declared in gvdb_amd.h but not exported: %s" % missing
    out = subprocess.run(["nm", "-D", "--defined-only", lib._name], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT\s+(gvdb_\w+)$", out, re.M))
    assert set(_declared()) <= exported
    assert exported <= set(_declared()), "exported but undeclared: %s" % sorted(exported - set(_declared()))


def test_python_binding_covers_the_header():
    import gvdb_amd
    assert sorted(gvdb_amd.exported_names()) == _declared()


def test_status_codes_and_enums_agree():
    import gvdb_amd as G
    d = _defines()
    assert d["GVDB_OK"] == G.OK
    for name in ("INVALID", "HIP", "STATE", "DEPTH", "NOBRICKS", "RANGE", "UNSUPPORTED", "OOM"):
        assert d["GVDB_ERR_" + name] == getattr(G, "ERR_" + name)
    assert d["GVDB_T_FLOAT"] == G.T_FLOAT and d["GVDB_T_HALF"] == G.T_HALF
    assert d["GVDB_SHADE_TRILINEAR"] == G.SHADE_TRILINEAR and d["GVDB_SHADE_VOLUME"] == G.SHADE_VOLUME


def test_node_record_layout(tmp_path):
    """gvdb_node as compiled from the header matches the 64-byte node the
    binding reads (src/gvdb_node.h:27-40 of the reference)."""
    import gvdb_amd as G
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gvdb_amd.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(gvdb_node),'
                   'offsetof(gvdb_node,pos), offsetof(gvdb_node,value), offsetof(gvdb_node,vrange),'
                   'offsetof(gvdb_node,parent), offsetof(gvdb_node,childlist), offsetof(gvdb_node,mask),'
                   'offsetof(gvdb_node,flags)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    dt = G.NODE_DTYPE
    assert vals == [dt.itemsize, dt.fields["mPos"][1], dt.fields["mValue"][1], dt.fields["mVRange"][1],
                    dt.fields["mParent"][1], dt.fields["mChildList"][1], dt.fields["mMask"][1], dt.fields["mFlags"][1]]
    assert dt.itemsize == 64


def test_null_handle_is_rejected(lib):
    import gvdb_amd as G
    assert lib.gvdb_destroy(None) == G.ERR_INVALID
    assert lib.gvdb_configure(None, 5, None, None) == G.ERR_INVALID
    assert lib.gvdb_render(None, 4, 0, 0) == G.ERR_INVALID
    assert lib.gvdb_last_error(None) == b"null handle"
    assert lib.gvdb_create(None, 0) == G.ERR_INVALID


def test_no_device_fails_loudly():
    """Without a HIP device the library refuses to create a volume; there is
    no CPU fallback."""
    import gvdb_amd as G
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a HIP device is present")
    h = ctypes.c_void_p()
    assert G.load().gvdb_create(ctypes.byref(h), 0) == G.ERR_HIP
    assert not h.value
    with pytest.raises(G.GVDBError):
        G.VolumeGVDB(0)


def test_product_does_not_import_the_oracle():
    """Nothing in the package or the C sources reaches the oracle."""
    pkg = os.path.join(ROOT, "gvdb-voxels_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower() or f == "build.py", os.path.join(dirpath, f)


DEMO = os.path.join(ROOT, "gvdb-voxels_amd", "lib", "facade_demo")


def test_cpp_facade_demo_fails_loudly_without_device(tmp_path):
    """The C++ facade (include/gvdb_amd.hpp) sample is built and reports the
    HIP status when no device is present."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a HIP device is present")
    pts = tmp_path / "p.f32"
    np.zeros((4, 3), np.float32).tofile(pts)
    r = subprocess.run([DEMO, str(pts), "4", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "gvdb error -2" in r.stderr


@pytest.mark.gpu
def test_cpp_facade_matches_python_binding(gpu, tmp_path):
    """The same pipeline through the C++ facade and through the Python binding
    gives the same node counts, entry count and image."""
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(20000, seed=41)
    f = tmp_path / "p.f32"
    pts.tofile(f)
    r = subprocess.run([DEMO, str(f), str(len(pts)), "2", "64"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tok = r.stdout.split()
    counts = [int(x) for x in tok[1:6]]
    entries, mean = int(tok[7]), float(tok[9])
    assert tok[10] == "batched_same" and tok[11] == "1"     # GridBatch in the C++ facade
    v, n = H.gpu_volume(gpu, pts, 2.0)
    assert counts == [len(v.pool(0, l)) for l in range(5)]
    assert entries == n
    sc = W.ea_scene(64, 64)
    v.apply_scene(sc)
    v.AddRenderBuf(0, 64, 64, 16)
    v.Render(gpu.SHADE_VOLUME, 0, 0)
    img = v.ReadRenderBuf(0, 64, 64, 16)
    assert abs(float(img.astype(np.float64).mean()) - mean) < 1e-6
    v.close()
