"""The frame loop (src/main.cpp:166-183, :228-261) and the display pass
(post-process-shader.frag:62-68) on the GPU, through the C-ABI: frames
accumulate exactly like rt_render_frames, a camera move restarts the
accumulation, and the display image is the gamma table applied to the
target."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CAM = [0.0, 3.0, 0.0]


@pytest.fixture(scope="module")
def soup(rt, gpu_ctx):
    hs = rt.scene_soup(1024, 1)
    return hs, gpu_ctx.upload(hs)


def test_frames_accumulate_like_render_frames(rt, gpu_ctx, soup):
    hs, scene = soup
    print("case: frame loop 3 frames 96x80", flush=True)
    loop = rt.FrameLoop(gpu_ctx, scene, 96, 80, vfov=60.0, bounces=4)
    try:
        for k in range(3):
            assert loop.current_sample == k
            loop.step(CAM, time=0.1 * k)
        tgt, acc = gpu_ctx.render_frames(scene, CAM, 96, 80, 3, vfov=60.0, bounces=4)
        torch.cuda.synchronize()
        assert np.array_equal(loop.image("target"), tgt.cpu().numpy())
        assert np.array_equal(loop.image("accum"), acc.cpu().numpy())
        assert loop.current_sample == 3
    finally:
        loop.close()


def test_move_restarts_accumulation(rt, gpu_ctx, soup):
    hs, scene = soup
    print("case: frame loop move after 2 frames", flush=True)
    loop = rt.FrameLoop(gpu_ctx, scene, 64, 48, vfov=60.0, bounces=4)
    try:
        loop.step(CAM)
        loop.step(CAM)
        moved = [0.5, 3.0, 0.2]
        loop.step(moved, moved=True)
        assert loop.current_sample == 1
        fresh = gpu_ctx.render(scene, moved, 64, 48, current_sample=0, vfov=60.0, bounces=4)
        torch.cuda.synchronize()
        assert np.array_equal(loop.image("target"), fresh.cpu().numpy())
    finally:
        loop.close()


def test_display_is_the_gamma_table_of_the_target(rt, gpu_ctx, soup):
    hs, scene = soup
    print("case: frame loop display 128x128", flush=True)
    loop = rt.FrameLoop(gpu_ctx, scene, 128, 128, vfov=60.0, bounces=4)
    try:
        loop.step(CAM)
        loop.display()
        tgt, disp = loop.image("target"), loop.image("display")
        lut = rt.gamma_table()
        assert np.array_equal(disp[..., :3], lut[tgt[..., :3]])
        assert (disp[..., 3] == 255).all()
    finally:
        loop.close()


def test_postprocess_on_any_image(rt, gpu_ctx):
    img = np.random.default_rng(5).integers(0, 256, (37, 53, 4), dtype=np.uint8)
    print("case: gamma pass 53x37", flush=True)
    out = gpu_ctx.postprocess_gamma(torch.from_numpy(img).cuda()).cpu().numpy()
    lut = rt.gamma_table()
    assert np.array_equal(out[..., :3], lut[img[..., :3]]) and (out[..., 3] == 255).all()
