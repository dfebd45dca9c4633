"""GPU parity of the render path (main / ray_color / scatter of
ray-trace-compute.comp) through the C-ABI, against the oracle's render on the
same seeded scene.  Bar: RGBA8 images within RMSE 1e-3 per channel (values
normalised to [0, 1]); the reference-order traversal is expected to match
exactly, and the ray counts are exact."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

CAM = [0.0, 3.0, 0.0]          # camPos (0,3,0) puts the eye at (0,0,3)


def oscene(hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def rmse(a, b):
    d = (a.astype(np.float64) - b.astype(np.float64)) / 255.0
    return float(np.sqrt((d * d).mean(axis=(0, 1))).max())


@pytest.fixture(scope="module")
def soup(rt, gpu_ctx):
    hs = rt.scene_soup(1024, 1)
    return hs, gpu_ctx.upload(hs)


def test_small_render_matches_oracle(rt, gpu_ctx, soup):
    """256x256, 1 spp, vfov 60, 4 bounces, currentSample 0."""
    hs, scene = soup
    print("case: render 256x256 packed", flush=True)
    img = gpu_ctx.render(scene, CAM, 256, 256, current_sample=0, vfov=60.0, bounces=4)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    want, rays = O.render(oscene(hs), CAM, 256, 256, current_sample=0, vfov=60.0, bounces=4, threads=8)
    e = rmse(got, want)
    ndiff = int((got != want).any(axis=2).sum())
    print(f"rmse {e:.2e}, differing pixels {ndiff}", flush=True)
    assert e <= 1e-3
    prim, sec = gpu_ctx.last_rays()
    assert prim + sec == rays
    assert (want[..., 3] == 255).all()


def test_small_render_reference_traversal_exact(rt, gpu_ctx, soup):
    import rtamd
    hs, scene = soup
    print("case: render reference traversal", flush=True)
    img = gpu_ctx.render(scene, CAM, 256, 256, vfov=60.0, bounces=4, traversal=rtamd.TRAVERSE_REFERENCE)
    torch.cuda.synchronize()
    want, _ = O.render(oscene(hs), CAM, 256, 256, vfov=60.0, bounces=4, threads=8)
    assert np.array_equal(img.cpu().numpy(), want)


def test_progressive_frames_match_oracle(rt, gpu_ctx, soup):
    """Four frames (currentSample 0..3) with the 8-bit accumulation copy."""
    hs, scene = soup
    print("case: render 4 progressive frames", flush=True)
    tgt, acc = gpu_ctx.render_frames(scene, CAM, 128, 96, 4, first_sample=0, vfov=60.0, bounces=4)
    torch.cuda.synchronize()
    want, rays = O.render_frames(oscene(hs), CAM, 128, 96, 4, vfov=60.0, bounces=4, threads=8)
    assert rmse(tgt.cpu().numpy(), want) <= 1e-3
    assert np.array_equal(tgt.cpu().numpy(), acc.cpu().numpy())
    prim, sec = gpu_ctx.last_rays()
    assert prim + sec == rays


def test_single_frame_with_accumulation_input(rt, gpu_ctx, soup):
    hs, scene = soup
    rng = np.random.default_rng(3)
    accum = rng.integers(0, 256, (64, 80, 4), dtype=np.uint8)
    print("case: render with accumulation input", flush=True)
    img = gpu_ctx.render(scene, CAM, 80, 64, current_sample=5, vfov=45.0, bounces=3,
                         accum=torch.from_numpy(accum).cuda())
    torch.cuda.synchronize()
    want, _ = O.render(oscene(hs), CAM, 80, 64, current_sample=5, vfov=45.0, bounces=3, accum=accum, threads=8)
    assert rmse(img.cpu().numpy(), want) <= 1e-3


def test_shader_defaults_and_odd_sizes(rt, gpu_ctx, soup):
    """vfov 30 and 2 bounces as the shader hard-codes them; an image size that
    is not a multiple of 32 is rendered whole."""
    hs, scene = soup
    print("case: render shader defaults 100x70", flush=True)
    img = gpu_ctx.render(scene, [0.5, 2.5, -0.3], 100, 70, current_sample=2)
    torch.cuda.synchronize()
    want, _ = O.render(oscene(hs), [0.5, 2.5, -0.3], 100, 70, current_sample=2, threads=8)
    assert rmse(img.cpu().numpy(), want) <= 1e-3


def test_render_with_spheres(rt, gpu_ctx, soup):
    hs, scene = soup
    print("case: render with spheres", flush=True)
    img = gpu_ctx.render(scene, CAM, 128, 128, vfov=60.0, bounces=4, spheres=1)
    torch.cuda.synchronize()
    want, _ = O.render(oscene(hs), CAM, 128, 128, vfov=60.0, bounces=4, spheres=1, traversal=O.TRAV_ORDERED,
                       threads=8)
    assert rmse(img.cpu().numpy(), want) <= 1e-3


def test_render_is_deterministic(rt, gpu_ctx, soup):
    hs, scene = soup
    a = gpu_ctx.render(scene, CAM, 192, 160, vfov=60.0, bounces=4).cpu().numpy()
    b = gpu_ctx.render(scene, CAM, 192, 160, vfov=60.0, bounces=4).cpu().numpy()
    assert np.array_equal(a, b)


def test_stage_timing_leaves_the_image_unchanged(rt, gpu_ctx, soup):
    """rt_render_set_timing only adds events: same image, same ray counts,
    and the stage times add up to the whole render."""
    hs, scene = soup
    print("case: timed render 160x128, 4 bounces", flush=True)
    plain = gpu_ctx.render(scene, CAM, 160, 128, vfov=60.0, bounces=4).cpu().numpy()
    st0 = gpu_ctx.last_render_stats()
    assert st0.timed == 0 and st0.total_ms == 0.0
    gpu_ctx.set_render_timing(True)
    try:
        timed = gpu_ctx.render(scene, CAM, 160, 128, vfov=60.0, bounces=4).cpu().numpy()
        st = gpu_ctx.last_render_stats()
    finally:
        gpu_ctx.set_render_timing(False)
    assert np.array_equal(plain, timed)
    assert st.timed == 1 and st.bounces == 4
    assert (st.primary_rays, st.secondary_rays) == (st0.primary_rays, st0.secondary_rays)
    assert st.primary_rays == 160 * 128 and st.secondary_rays > 0
    assert st.trace_primary_ms > 0 and st.trace_secondary_ms > 0 and st.shade_ms >= 0
    assert st.trace_primary_ms + st.trace_secondary_ms <= st.total_ms * (1 + 1e-5)


def test_render_with_hit_scene_sphere_rule(rt, gpu_ctx, soup):
    """RT_SPHERES_SCENE in a render: a sphere hit overrides a nearer triangle
    (hit_scene, ray-trace-compute.comp:222-247); shader-order traversal,
    bit-exact against the oracle."""
    hs, scene = soup
    print("case: render with hit_scene's sphere rule", flush=True)
    img = gpu_ctx.render(scene, CAM, 128, 128, vfov=60.0, bounces=4, spheres=2,
                         traversal=rt.TRAVERSE_REFERENCE).cpu().numpy()
    want, _ = O.render(oscene(hs), CAM, 128, 128, vfov=60.0, bounces=4, spheres=2, threads=8)
    assert np.array_equal(img, want)
    near, _ = O.render(oscene(hs), CAM, 128, 128, vfov=60.0, bounces=4, spheres=1, threads=8)
    assert not np.array_equal(want, near)            # the two rules differ on this scene


def test_bench_render_against_the_oracle_on_sampled_rows(rt, gpu_ctx):
    """The bench render -- 4,194,304-triangle clustered soup, 1920x1080,
    4 progressive frames, 8 bounces, vfov 60, eye (0,0,10) -- through the
    default path (sorted secondary bounces, XCD segments), against the
    oracle's ordered traversal on every 72nd pixel row: RMSE <= 1e-3."""
    hs = rt.scene_clusters(1 << 22, 3)
    scene = gpu_ctx.upload(hs)
    print("case: bench render 1920x1080x4, 8 bounces, oracle on every 72nd row", flush=True)
    tgt, _ = gpu_ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8)
    torch.cuda.synchronize()
    got = tgt.cpu().numpy()[::72]
    want, _ = O.render_frames(oscene(hs), [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8,
                              traversal=O.TRAV_ORDERED, row_step=72, threads=16)
    want = want[::72]
    assert got[..., :3].mean() > 1.0
    assert rmse(got, want) <= 1e-3


def test_render_record_counts(rt, gpu_ctx):
    """A counted render gives the image of an uncounted one, and with one
    bounce its record totals equal the counting query's over the same camera
    rays (generated on the host with the shader's camera arithmetic)."""
    hs = rt.scene_clusters(200000, 9, threads=16)
    sc = gpu_ctx.upload(hs)
    W, H, cam = 320, 240, [0.0, 10.0, 0.0]
    print("case: counted render, 200k-triangle clustered soup, 320x240", flush=True)
    img = gpu_ctx.render(sc, cam, W, H, vfov=60.0, bounces=4).cpu().numpy()
    gpu_ctx.set_render_counting(True)
    try:
        counted = gpu_ctx.render(sc, cam, W, H, vfov=60.0, bounces=4).cpu().numpy()
        nv4, nt4 = gpu_ctx.last_render_counts()
        gpu_ctx.render(sc, cam, W, H, vfov=60.0, bounces=1)
        nv1, nt1 = gpu_ctx.last_render_counts()
    finally:
        gpu_ctx.set_render_counting(False)
    assert np.array_equal(img, counted)
    rays = rt.rays_camera_tiles(cam, 60.0, W, H, 1, 0, W * H)
    _, nv, tt = gpu_ctx.closest_hit_stats(sc, rt.rays_to_device(rays))
    assert nv1 == int(nv.to(torch.int64).sum().item()) and nt1 == int(tt.to(torch.int64).sum().item())
    assert nv4 > nv1 > 0 and nt4 >= nt1
    with pytest.raises(rt.RtError):
        gpu_ctx.render(sc, cam, W, H, vfov=60.0, bounces=1)
        gpu_ctx.last_render_counts()                     # the last render was not counted
    sc.release()
