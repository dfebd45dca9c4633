"""The reference's own scene (RtScene.h:44-101, 1,422 triangles from its
doge OBJ models, kept as tests/golden/doge_triangles.npy) rendered on the
GPU with the shader's own settings -- vfov 30, two bounces, the camera of
main.cpp:37 -- against the oracle's render of the same scene.  Bar: RMSE
<= 1e-3 per channel; the shader-order traversal with its 16-entry stack
matches exactly."""
import os

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "doge_triangles.npy")
CAM = [1.8, 8.6, 1.1]             # Camera camera(glm::vec3(1.8f, 8.6f, 1.1f)), main.cpp:37


def doge(rt):
    tris = np.frombuffer(bytearray(np.load(GOLDEN).tobytes()), rt.TRI_DT).copy()
    mats = np.zeros(6, rt.MAT_DT)                  # RtScene.h:49-62
    mats["type"] = [rt.RT_MAT_LAMBERTIAN] * 3 + [rt.RT_MAT_LIGHT, rt.RT_MAT_METAL, rt.RT_MAT_GLASS]
    mats["albedo"] = np.array([[.3, .3, .3], [.9, .1, .1], [.1, .9, .1], [2, 2, 2], [1, 1, 1], [1, 1, 1]], np.float32)
    spheres = np.zeros(1, rt.SPHERE_DT)            # RtScene.h:97
    spheres["s"] = np.array([[0.6, 1.0, -1.0, 0.6]], np.float32)
    spheres["materialIndex"] = 5
    return rt.HostScene(tris, mats, rt.bvh_build(tris, seed=1), rt.scene_lights(tris, mats), spheres)


@pytest.fixture(scope="module")
def scene(rt, gpu_ctx):
    hs = doge(rt)
    return hs, gpu_ctx.upload(hs)


def oscene(hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def rmse(a, b):
    d = (a.astype(np.float64) - b.astype(np.float64)) / 255.0
    return float(np.sqrt((d * d).mean(axis=(0, 1))).max())


def test_doge_render_matches_oracle(rt, gpu_ctx, scene):
    hs, sc = scene
    assert sc.info.num_triangles == 1422
    print("case: doge 320x240, shader defaults", flush=True)
    img = gpu_ctx.render(sc, CAM, 320, 240, current_sample=0).cpu().numpy()
    want, rays = O.render(oscene(hs), CAM, 320, 240, current_sample=0, threads=8)
    assert rmse(img, want) <= 1e-3
    assert sum(gpu_ctx.last_rays()) == rays
    assert img[..., :3].mean() > 1.0               # the light reaches the camera


def test_doge_shader_order_traversal_exact(rt, gpu_ctx, scene):
    hs, sc = scene
    print("case: doge reference traversal, 16-entry stack", flush=True)
    img = gpu_ctx.render(sc, CAM, 160, 120, current_sample=3, max_stack=16,
                         traversal=rt.TRAVERSE_REFERENCE).cpu().numpy()
    want, _ = O.render(oscene(hs), CAM, 160, 120, current_sample=3, max_stack=16, threads=8)
    assert np.array_equal(img, want)


def test_doge_progressive_frames(rt, gpu_ctx, scene):
    hs, sc = scene
    print("case: doge 8 progressive frames", flush=True)
    tgt, _ = gpu_ctx.render_frames(sc, CAM, 128, 96, 8)
    torch.cuda.synchronize()
    want, _ = O.render_frames(oscene(hs), CAM, 128, 96, 8, threads=8)
    assert rmse(tgt.cpu().numpy(), want) <= 1e-3
