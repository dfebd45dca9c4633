"""Spheres on the GPU (the pass after the triangle traversal): hit_scene's
own rule (ray-trace-compute.comp:222-247) and the nearest-hit rule, with a
plain loop for a few spheres and the sphere tree for many, against the
oracle's loop over every sphere.  Bar: with the shader-order traversal the
hits are bit-identical; the sphere tree returns exactly the loop's sphere
(smallest accepted root, highest index between equal roots), grazing rays
and duplicate spheres included."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def oscene(hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def run(rt, ctx, scene, rays, **kw):
    h = ctx.closest_hit(scene, rt.rays_to_device(rays), **kw)
    torch.cuda.synchronize()
    return rt.hits_from_device(h)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bit_equal(got, want):
    bad = np.nonzero(got["id"] != want["id"])[0]
    assert len(bad) == 0, (bad[:8], got[bad[:4]], want[bad[:4]])
    for f in ("t", "u", "v"):
        assert np.array_equal(bits(got[f]), bits(want[f])), f


def many_spheres(rt, n, seed):
    rng = np.random.default_rng(seed)
    sp = np.zeros(n, rt.SPHERE_DT)
    c = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(0.005), np.log(0.1), n)).astype(np.float32)
    sp["s"] = np.concatenate([c, r[:, None]], axis=1)
    sp["materialIndex"] = rng.integers(0, 4, n)
    sp[n - 40:] = sp[:40]                       # duplicates: equal roots, the higher index must win
    return sp


def grazing_rays(rt, sp, n, seed):
    """Rays from radius-3 origins aimed just inside / outside sphere silhouettes."""
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, rt.RAY_DT)
    j = rng.integers(0, len(sp), n)
    c = sp["s"][j, :3].astype(np.float64)
    r = sp["s"][j, 3].astype(np.float64)
    o = rng.normal(size=(n, 3))
    o = 3.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    w = c - o
    u = np.cross(w, rng.normal(size=(n, 3)))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    k = 1.0 + rng.uniform(-1e-5, 1e-5, n)
    d = c + (r * k)[:, None] * u - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"], rays["d"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 0.001, 10000.0
    return rays


@pytest.fixture(scope="module")
def soup(rt, gpu_ctx):
    hs = rt.scene_soup(1024, 1)
    return hs, gpu_ctx.upload(hs)


@pytest.fixture(scope="module")
def crowd(rt, gpu_ctx):
    base = rt.scene_soup(1024, 1)
    hs = rt.HostScene(base.triangles, base.materials, base.nodes, base.lights, many_spheres(rt, 3000, 12))
    return hs, gpu_ctx.upload(hs)


def mixed(rt, n, seed):
    return np.concatenate([rt.rays_shell(n // 2, seed, 3.0, 0.5), rt.rays_box(n - n // 2, seed + 1, 1.5)])


def test_scene_rule_two_spheres(rt, gpu_ctx, soup):
    hs, scene = soup
    rays = mixed(rt, 32768, 21)
    print("case: soup, hit_scene sphere rule, reference traversal", flush=True)
    want = O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, spheres=2, threads=8)
    got = run(rt, gpu_ctx, scene, rays, spheres=rt.SPHERES_SCENE, traversal=rt.TRAVERSE_REFERENCE)
    assert_bit_equal(got, want)
    assert (want["id"] >= 1024).sum() > 100          # spheres do override here


@pytest.mark.parametrize("rule", [1, 2])
def test_sphere_tree_equals_the_loop(rt, gpu_ctx, crowd, rule):
    hs, scene = crowd
    sc = oscene(hs)
    rays = np.concatenate([mixed(rt, 20000, 31), grazing_rays(rt, hs.spheres, 20000, 32)])
    print(f"case: 3000 spheres, rule {rule}, reference traversal", flush=True)
    want = O.closest_hit(sc, rays, O.TRAV_REFERENCE, spheres=rule, threads=8)
    got = run(rt, gpu_ctx, scene, rays, spheres=rule, traversal=rt.TRAVERSE_REFERENCE)
    assert_bit_equal(got, want)
    hit_sphere = want["id"] >= 1024
    assert hit_sphere.sum() > 5000
    dup = (want["id"] >= 1024 + 3000 - 40)
    assert dup.sum() > 0                              # duplicates were hit, and the later copy won


def test_sphere_tree_with_packed_traversal(rt, gpu_ctx, crowd):
    hs, scene = crowd
    rays = mixed(rt, 40000, 41)
    print("case: 3000 spheres, nearest rule, packed traversal", flush=True)
    want = O.closest_hit(oscene(hs), rays, O.TRAV_ORDERED, spheres=1, threads=8)
    got = run(rt, gpu_ctx, scene, rays, spheres=rt.SPHERES_NEAREST)
    sph = want["id"] >= 1024
    assert np.array_equal(got["id"][sph], want["id"][sph])
    for f in ("t", "u", "v"):
        assert np.array_equal(bits(got[f][sph]), bits(want[f][sph])), f
    assert (got["id"] == want["id"]).mean() > 0.9999


def test_render_with_many_spheres(rt, gpu_ctx, crowd):
    hs, scene = crowd
    print("case: render 128x96 with 3000 spheres", flush=True)
    img = gpu_ctx.render(scene, [0.0, 3.0, 0.0], 128, 96, vfov=60.0, bounces=4, spheres=1,
                         traversal=rt.TRAVERSE_REFERENCE).cpu().numpy()
    want, _ = O.render(oscene(hs), [0.0, 3.0, 0.0], 128, 96, vfov=60.0, bounces=4, spheres=1, threads=8)
    assert np.array_equal(img, want)


def test_unknown_sphere_rule_rejected(rt, gpu_ctx, soup):
    hs, scene = soup
    with pytest.raises(rt.RtError):
        gpu_ctx.closest_hit(scene, rt.rays_to_device(mixed(rt, 64, 1)), spheres=3)
