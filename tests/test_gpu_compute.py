"""Compute operators on the GPU (Compute, src/gvdb_volume_gvdb.cpp:4527-4552;
kernels/cuda_gvdb_operators.cuh:148-455 of the reference) against the
oracle's FP32 restatement started from the same atlas: smooth (with the
apron refreshed between iterations, as the reference's point samples run
it), noise, grow and fill are bit-exact; the sine fill, whose sinf differs
between math libraries, within 2e-6."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu


def _pair(gpu):
    import oracle as O
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(20000, seed=91)
    v, _ = H.gpu_volume(gpu, pts, 2.0)
    o, _ = H.oracle_volume(O, pts, 2.0, precision=32)
    assert v.atlas_info() == o.atlas_info()
    o.set_atlas(v.atlas(), 32)
    return v, o


@pytest.mark.parametrize("effect,iters,params,apron,skip", [
    (113, 3, (2.0, 0.0, 0.0), True, True),       # the point samples' smoothing (8 iterations there)
    (113, 1, (0.5, 0.1, 0.0), False, True),
    (114, 1, (0.3, 0.0, 0.0), False, True),
    (115, 2, (0.05, 0.0, 0.0), True, True),
    (110, 1, (1.5, 0.0, 0.0), False, False),
    (110, 1, (1.5, 0.0, 0.0), False, True),
])
def test_compute_matches_oracle(gpu, effect, iters, params, apron, skip):
    v, o = _pair(gpu)
    v.Compute(effect, 0, iters, params, apron, skip, 0.0)
    v.synchronize()
    assert o.compute(effect, iters, params, apron, skip, 0.0, 32) == 0
    a, b = v.atlas(), o.atlas(32)
    mask = H.used_brick_mask(b.shape, o.atlas_map(), o.atlas_info()["cnt"])
    if effect == 110:
        mask = np.ones_like(mask)
    assert np.array_equal(a[mask], b[mask]), "%d voxels differ" % int((a[mask] != b[mask]).sum())
    v.close()


def test_compute_sine_fill(gpu):
    v, o = _pair(gpu)
    v.Compute(110, 0, 1, (0.0, 0.0, -1.0), False, False, 0.0)
    v.synchronize()
    assert o.compute(110, 1, (0.0, 0.0, -1.0), False, False, 0.0, 32) == 0
    assert np.abs(v.atlas() - o.atlas(32)).max() <= 2e-6
    v.close()


def test_compute_refuses_uchar_effects(gpu):
    from gvdb_amd import GVDBError
    v, _ = H.gpu_volume(gpu, (np.random.default_rng(3).random((500, 3)) * 40 + 70).astype(np.float32), 2.0)
    for eff in (111, 112, 116, 117, 1):
        with pytest.raises(GVDBError) as e:
            v.Compute(eff, 0, 1, (0, 0, 0), False, True)
        assert e.value.code == -7
    v.close()
