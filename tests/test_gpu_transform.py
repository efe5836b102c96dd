"""SetTransform in the march (PrepareRender, gvdb_volume_gvdb.cpp:4254-4306;
SetTransform, :5770-5793): the camera rays and the light go into index space
through the inverse of T R S PT.  Closed forms: moving the grid and the camera
(and light) by the same translation, or turning both by the same rotation
about y, leaves the image; a scale s with the camera s times as far (fov
fixed) too.  And the GPU march with a general transform against the oracle's
FP32 march over the same atlas and scene block."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

C = np.array([96.0, 90.0, 100.0])


def _grid(gpu):
    from gvdb_amd import workloads as W
    pts = W.sphere_shell(20000, seed=101, radius=40.0, centre=tuple(C))
    v, _ = H.gpu_volume(gpu, pts, 2.0)
    return v


def _render(v, shading, to, angs, dist, light):
    from gvdb_amd import workloads as W
    w = 128
    sc = (W.ea_scene if shading == 7 else W.iso_scene)(w, w, centre=tuple(float(x) for x in to), dist=float(dist))
    sc["cam"]["angs"] = angs
    sc["light"] = tuple(float(x) for x in light)
    v.apply_scene(sc)
    v.AddRenderBuf(0, w, w, 16)
    v.Render(shading, 0, 0)
    return v.ReadRenderBuf(0, w, w, 16), sc


@pytest.mark.parametrize("shading", [7, 4])
def test_translation_and_rotation_move_the_camera_with_the_grid(gpu, shading):
    v = _grid(gpu)
    L = C + np.array([300.0, 400.0, 200.0])
    ref, _ = _render(v, shading, C, (30.0, 20.0, 0.0), 320.0, L)
    t = np.array([50.0, -20.0, 35.0])
    v.SetTransform((0, 0, 0), (1, 1, 1), (0, 0, 0), tuple(t))
    img, _ = _render(v, shading, C + t, (30.0, 20.0, 0.0), 320.0, L + t)
    assert H.rmse(img, ref) <= 1e-3
    th = 40.0
    c, s = np.cos(np.radians(th)), np.sin(np.radians(th))
    Ry = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    v.SetTransform((0, 0, 0), (1, 1, 1), (0, th, 0), (0, 0, 0))
    img, _ = _render(v, shading, Ry @ C, (30.0 + th, 20.0, 0.0), 320.0, Ry @ L)
    assert H.rmse(img, ref) <= 1e-3
    v.SetTransform((0, 0, 0), (2, 2, 2), (0, 0, 0), (0, 0, 0))
    img, _ = _render(v, shading, 2 * C, (30.0, 20.0, 0.0), 640.0, 2 * L)
    assert H.rmse(img, ref) <= 1e-3
    v.close()


@pytest.mark.parametrize("shading", [7, 4])
def test_general_transform_against_the_oracle_march(gpu, shading):
    import oracle as O
    v = _grid(gpu)
    v.SetTransform((-96.0, -90.0, -100.0), (1.5, 0.75, 1.25), (20.0, -35.0, 60.0), (10.0, 5.0, -8.0))
    img, sc = _render(v, shading, (10.0, 5.0, -8.0), (30.0, 20.0, 0.0), 260.0, (300.0, 400.0, 200.0))
    assert np.count_nonzero(np.abs(img - np.asarray(sc["backclr"], np.float32)).max(-1) > 1e-3) > 500
    from gvdb_amd import workloads as W
    o, _ = H.oracle_volume(O, W.sphere_shell(20000, seed=101, radius=40.0, centre=tuple(C)), 2.0, precision=32)
    o.set_atlas(v.atlas(), 32)
    ref = o.render(v.scene_block(128, 128), O.linear_tf(*sc["tf"]), shading, 32)
    assert H.rmse(img, ref) <= 1e-3
    v.close()
