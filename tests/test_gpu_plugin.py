"""User kernels over a volume (ComputeKernel, RenderKernel;
src/gvdb_volume_gvdb.cpp:4499-4525 and :4309-4333 of the reference): the
sample code object built with the library is loaded through the C ABI, a
compute kernel rewrites every brick interior (2 v + 1) and leaves the aprons,
and a render kernel sees the launch shape and the view of the volume."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _interior_mask(shape, cnt, br=10):
    m = np.zeros(shape, bool)
    for sz in range(shape[0] // br):
        for sy in range(cnt[1]):
            for sx in range(cnt[0]):
                m[sz * br + 1:sz * br + br - 1, sy * br + 1:sy * br + br - 1, sx * br + 1:sx * br + br - 1] = True
    return m


@pytest.fixture
def kernels(gpu):
    mod = gpu.module_load(gpu.user_kernels_path())
    yield gpu.module_function(mod, "user_scale"), gpu.module_function(mod, "user_pixels")
    gpu.module_unload(mod)


def test_compute_kernel_scales_brick_interiors(gpu, kernels):
    from gvdb_amd import workloads as W
    v, _ = H.gpu_volume(gpu, W.sphere_shell(20000, seed=95), 2.0)
    a0 = v.atlas()
    v.ComputeKernel(kernels[0], 0, False, True)
    a1 = v.atlas()
    m = _interior_mask(a0.shape, v.atlas_info()["cnt"])
    assert np.array_equal(a1[m], np.float32(2) * a0[m] + np.float32(1))
    assert np.array_equal(a1[~m], a0[~m])
    # with the apron refreshed: the aprons match a separate UpdateApron
    v.ComputeKernel(kernels[0], 0, True, True)
    a2 = v.atlas()
    v.UpdateApron(0, 0.0)
    assert np.array_equal(v.atlas(), a2)
    v.close()


def test_render_kernel_launch_and_view(gpu, kernels):
    from gvdb_amd import workloads as W
    v, _ = H.gpu_volume(gpu, W.sphere_shell(20000, seed=96), 2.0)
    sc = W.iso_scene(100, 70, centre=(256.0, 256.0, 256.0), dist=600.0)
    v.apply_scene(sc)
    v.AddRenderBuf(0, 100, 70, 16)
    v.RenderKernel(kernels[1], 0, 0)
    img = v.ReadRenderBuf(0, 100, 70, 16).reshape(70, 100, 4)
    yy, xx = np.mgrid[0:70, 0:100]
    assert np.array_equal(img[..., 0], xx) and np.array_equal(img[..., 1], yy)
    assert (img[..., 2] == 0).all() and (img[..., 3] == v.atlas_info()["used"]).all()
    v.close()


def test_module_errors(gpu):
    from gvdb_amd import GVDBError
    with pytest.raises(GVDBError):
        gpu.module_load("/nonexistent/kernels.hsaco")
    mod = gpu.module_load(gpu.user_kernels_path())
    with pytest.raises(GVDBError):
        gpu.module_function(mod, "no_such_kernel")
    gpu.module_unload(mod)
