"""Known-answer tests that pin the oracle (oracle/rt_oracle.c) to values the
reference's mathematics fixes, independently of the oracle's own code.

Each check is chosen so a plausible restatement mistake fails: a swapped
cross-product operand, swapped barycentrics, a wrong sign, an inclusive/open
interval mix-up, a wrong RNG constant or shift, or the wrong draw order.
"""
import math

import numpy as np
import pytest

F = np.float32


def tri(v0, v1, v2, mat=0):
    from oracle import TRI_DT
    t = np.zeros(1, TRI_DT)
    t["v0"], t["v1"], t["v2"], t["materialIndex"] = v0, v1, v2, mat
    return t


# --------------------------------------------------------------------- RNG
# random.glsl:4-22.  Values derived from the integer formula by a separate
# computation (not the oracle), for (x, y, currentSample).
RNG_KAT = {
    (0, 0, 0): ([0x108EF29B, 0x00033628, 0xDCCC2102], [0x3D847795, 0x384D8A00, 0x3F5CCC21]),
    (1, 0, 0): ([0x9F21590F, 0x44E67BC7, 0x60B37254], [0x3F1F2159, 0x3E89CCF8, 0x3EC166E5]),
    (37, 200, 0): ([0xECE8A6C2, 0xD50B0B38, 0x29AA1E89], [0x3F6CE8A7, 0x3F550B0B, 0x3E26A87A]),
    (37, 200, 3): ([0xD6E355BE, 0x3727AD87, 0x73888EFD], [0x3F56E356, 0x3E5C9EB6, 0x3EE7111E]),
}


@pytest.mark.parametrize("key", sorted(RNG_KAT))
def test_rng_known_answers(orc, key):
    f, w = orc.rng_stream(*key, 3)
    words, bits = RNG_KAT[key]
    assert [int(v) for v in w] == words
    assert [int(v) for v in f.view(np.uint32)] == bits


def test_rng_seed_and_shift(orc):
    # seed = (600x + y) * (currentSample + 1) with uint wrap-around
    assert orc.rng_seed(37, 200, 3) == ((600 * 37 + 200) * 4) & 0xFFFFFFFF
    assert orc.rng_seed(10_000_000, 5, 1000) == ((600 * 10_000_000 + 5) * 1001) & 0xFFFFFFFF
    # seed 0 steps to state 1, so (0, 1, 0) is (0, 0, 0) shifted by one draw
    a, _ = orc.rng_stream(0, 0, 0, 6)
    b, _ = orc.rng_stream(0, 1, 0, 5)
    assert np.array_equal(a[1:], b)


def test_rng_float_is_exact_quotient(orc):
    # float(word) / 4294967295.0f: the divisor rounds to 2^32, so the result
    # is float(word) * 2^-32 exactly and lies in [0, 1] inclusive.
    f, w = orc.rng_stream(123, 456, 7, 4096)
    expect = (w.astype(np.float32) * F(2.0 ** -32)).astype(np.float32)
    assert np.array_equal(f, expect)
    assert f.min() >= 0.0 and f.max() <= 1.0


def test_glibc_rand_restatement_matches_libc():
    import ctypes
    import subprocess
    import sys
    from oracle.bvh_ref import GlibcRandom
    # a fresh interpreter so nothing has consumed rand() yet
    code = "import ctypes;L=ctypes.CDLL('libc.so.6');print(' '.join(str(L.rand()) for _ in range(2000)))"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout.split()
    g = GlibcRandom(1)
    assert [g.next() for _ in range(2000)] == [int(v) for v in out]
    L = ctypes.CDLL("libc.so.6")
    L.srand(12345)
    g = GlibcRandom(12345)
    assert [g.next() for _ in range(500)] == [L.rand() for _ in range(500)]


# ---------------------------------------------------------------- AABB test
def test_aabb_slab_closed_forms(orc):
    # tMin = (1, .5, .25), tMax = (2, 1, .5): tNear = 1 > tFar = .5 -> miss
    tn, tf = orc.intersect_aabb([0, 0, 0], [1, 2, 4], [1, 1, 1], [2, 2, 2])
    assert (tn, tf) == (1.0, 0.5)
    # zero direction components divide to +-inf and drop out of the slab test
    tn, tf = orc.intersect_aabb([0, 0.5, 0.5], [1, 0, 0], [1, 0, 0], [2, 1, 1])
    assert (tn, tf) == (1.0, 2.0)
    # box behind the origin: still tNear <= tFar (the shader's line test visits it)
    tn, tf = orc.intersect_aabb([0, 0.5, 0.5], [1, 0, 0], [-3, 0, 0], [-2, 1, 1])
    assert (tn, tf) == (-3.0, -2.0) and tn <= tf
    # origin inside: tNear < 0 < tFar
    tn, tf = orc.intersect_aabb([0.5, 0.5, 0.5], [0, 0, -1], [0, 0, 0], [1, 1, 1])
    assert tn == -0.5 and tf == 0.5


# ------------------------------------------------------------ triangle test
UNIT_TRI = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


def test_tri_intersect_barycentrics_and_sign(orc):
    tuv, n = orc.tri_intersect([0.25, 0.5, 1.0], [0, 0, -1], tri(*UNIT_TRI))
    # u weights v1 (the x axis here), v weights v2 (y); dyadic values are exact
    assert tuple(tuv) == (1.0, 0.25, 0.5)
    # n = cross(b, a) with a = v0 - v1, b = v2 - v0
    assert tuple(n) == (0.0, 0.0, 1.0)


@pytest.mark.parametrize("o,ok", [
    ([0.0, 0.5, 1.0], True),     # u = 0: edges are inclusive
    ([0.5, 0.5, 1.0], True),     # u + v = 1 exactly
    ([0.5, 0.0, 1.0], True),     # v = 0
    ([0.6, 0.5, 1.0], False),    # u + v > 1
    ([-0.1, 0.5, 1.0], False),   # u < 0
])
def test_hit_triangle_edges(orc, o, ok):
    hit = orc.hit_triangle(tri(*UNIT_TRI), o, [0, 0, -1])[0]
    assert hit is ok


def test_hit_triangle_open_t_interval(orc):
    t = tri(*UNIT_TRI)
    # t == 0.001 exactly is rejected (t > tMin), slightly above is accepted
    assert orc.hit_triangle(t, [0.25, 0.5, 0.001], [0, 0, -1], tmin=0.001)[0] is False
    assert orc.hit_triangle(t, [0.25, 0.5, 0.001], [0, 0, -1], tmin=0.0009)[0] is True
    # t == tMax is rejected (t < tMax)
    assert orc.hit_triangle(t, [0.25, 0.5, 2.0], [0, 0, -1], tmax=2.0)[0] is False
    # parallel ray: 1/0 -> inf, t is not finite and fails the range test
    assert orc.hit_triangle(t, [0.2, 0.2, 1.0], [1, 0, 0])[0] is False


def test_hit_triangle_record_front_and_back(orc):
    t = tri(*UNIT_TRI, mat=7)
    ok, p, nrm, tt, back, mat = orc.hit_triangle(t, [0.25, 0.5, 1.0], [0, 0, -1])
    assert ok and back == 0 and mat == 7 and tt == 1.0
    assert tuple(nrm) == (0.0, 0.0, 1.0)
    assert tuple(p) == (0.25, 0.5, float(F(1e-4)))          # p offset along the normal
    ok, p, nrm, tt, back, mat = orc.hit_triangle(t, [0.25, 0.5, -1.0], [0, 0, 1])
    assert ok and back == 1
    assert tuple(nrm) == (0.0, 0.0, -1.0)                    # flipped to face the ray
    assert tuple(p) == (0.25, 0.5, float(-F(1e-4)))


# --------------------------------------------------------------- sphere test
def sph(c, r, mat=0):
    from oracle import SPHERE_DT
    s = np.zeros(1, SPHERE_DT)
    s["s"] = [c[0], c[1], c[2], r]
    s["materialIndex"] = mat
    return s


def test_hit_sphere_closed_forms(orc):
    ok, p, n, t, back, mat = orc.hit_sphere(sph([0, 0, 0], 1.0, 3), [0, 0, 5], [0, 0, -1])
    assert ok and t == 4.0 and back == 0 and mat == 3
    assert tuple(p) == (0.0, 0.0, 1.0) and tuple(n) == (0.0, 0.0, 1.0)   # p is not offset
    # unnormalised direction: a = 4, half_b = -10, c = 24 -> root 2
    ok, p, n, t, back, _ = orc.hit_sphere(sph([0, 0, 0], 1.0), [0, 0, 5], [0, 0, -2])
    assert ok and t == 2.0 and tuple(p) == (0.0, 0.0, 1.0)
    # miss
    assert orc.hit_sphere(sph([0, 0, 0], 1.0), [0, 2, 5], [0, 0, -1])[0] is False


def test_hit_sphere_inside_normal_points_inward(orc):
    ok, p, n, t, back, _ = orc.hit_sphere(sph([0, 0, 0], 1.0), [0, 0, 0], [0, 0, -1])
    assert ok and back == 1 and t == 1.0
    assert tuple(p) == (0.0, 0.0, -1.0)
    # normal = (1 - 2*back) * (p - c) / r = -(0, 0, -1) = (0, 0, 1)
    assert n[2] == 1.0


# ------------------------------------------------------------------ scatter
def scatter_scene(orc):
    from oracle import MAT_DT, LIGHT_DT
    mats = np.zeros(4, MAT_DT)
    mats["type"] = [1, 2, 3, 0]
    mats["albedo"] = [[0.5, 0.25, 0.125], [0.9, 0.8, 0.7], [0.3, 0.3, 0.3], [4, 4, 4]]
    light = tri([-1, 2, -1], [1, 2, -1], [-1, 2, 1], mat=3)
    tris = np.concatenate([tri(*UNIT_TRI), light])
    lights = np.zeros(1, LIGHT_DT)
    lights["triangleIndex"] = 1
    lights["area"] = 2.0
    nodes = np.zeros(1, orc.NODE_DT)
    return orc.Scene(tris, mats, nodes, lights)


def step(s):
    return (s * 747796405 + 1) & 0xFFFFFFFF


def draws(state, n):
    out = []
    for _ in range(n):
        state = step(state)
        w = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
        w = ((w >> 22) ^ w) & 0xFFFFFFFF
        out.append(float(np.float32(w) / np.float32(4294967295.0)))
    return out


def test_metal_reflects_and_draws_once(orc):
    sc = scatter_scene(orc)
    e, alb, d, st = orc.scatter(sc, [1, -1, 0], [0, 0, 0], [0, 1, 0], 0, 1, 99)
    assert not e and tuple(d) == (1.0, 1.0, 0.0)              # I - 2 dot(N,I) N
    assert np.allclose(alb, [0.9, 0.8, 0.7]) and st == step(99)


def test_glass_refracts_by_snell(orc):
    sc = scatter_scene(orc)
    I = np.array([1, -1, 0], np.float32)
    e, alb, d, st = orc.scatter(sc, I, [0, 0, 0], [0, 1, 0], 0, 2, 5)
    s_t = (1 / 1.5) * math.sin(math.pi / 4)
    assert np.allclose(d, [s_t, -math.sqrt(1 - s_t * s_t), 0], atol=2e-7)
    assert tuple(alb) == (1.0, 1.0, 1.0) and st == step(5)    # glass forces albedo 1


def test_glass_total_internal_reflection(orc):
    sc = scatter_scene(orc)
    a = math.radians(60)
    I = np.array([math.sin(a), -math.cos(a), 0], np.float32)
    # back face: ratio = 1.5, 1.5 * sin 60 > 1 -> reflect
    e, alb, d, st = orc.scatter(sc, I, [0, 0, 0], [0, 1, 0], 1, 2, 5)
    assert np.allclose(d, [math.sin(a), math.cos(a), 0], atol=2e-7)


def test_light_emits_and_draws_once(orc):
    sc = scatter_scene(orc)
    e, alb, d, st = orc.scatter(sc, [0, -1, 0], [0, 0, 0], [0, 1, 0], 0, 3, 42)
    assert e and tuple(alb) == (4.0, 4.0, 4.0) and st == step(42)


@pytest.mark.parametrize("state", [1, 2, 3, 17, 1000, 123456789, 0xDEADBEEF])
def test_lambertian_draws_and_light_sampling(orc, state):
    sc = scatter_scene(orc)
    p = np.array([0.1, 0.0, 0.2], np.float32)
    n = np.array([0, 1, 0], np.float32)
    e, alb, d, st = orc.scatter(sc, [0, -1, 0], p, n, 0, 0, state)
    r = draws(state, 6)
    assert not e and np.allclose(alb, [0.5, 0.25, 0.125])
    assert abs(np.linalg.norm(d) - 1.0) < 1e-6
    s = state
    if r[2] < 0.5:
        # light branch: index floor(1 * r3) = 0, point on the parallelogram
        for _ in range(6):
            s = step(s)
        assert st == s
        v0, v1, v2 = np.array([-1, 2, -1.0]), np.array([1, 2, -1.0]), np.array([-1, 2, 1.0])
        pt = v0 + r[4] * (v1 - v0) + r[5] * (v2 - v0)
        want = (pt - p) / np.linalg.norm(pt - p)
        if abs(want[1]) >= 0.001:
            assert np.allclose(d, want, atol=2e-6)
    else:
        for _ in range(3):
            s = step(s)
        assert st == s
        # cosine direction about the normal: z = sqrt(1 - r2) is the normal component
        assert abs(float(d @ n) - math.sqrt(1 - r[1])) < 2e-6


# ------------------------------------------------------------ accumulation
def test_accumulate_first_sample_quantises(orc):
    out = orc.accumulate([0.5, 1.2, -1.0], [9, 9, 9, 9], 0)
    # 0.5 * 255 = 127.5 rounds to even; > 1 clamps to 255; < 0 clamps to 0; alpha 1
    assert list(out) == [128, 255, 0, 255]


def test_accumulate_running_average(orc):
    # currentSample 1: (c + acc/255) / 2
    out = orc.accumulate([0.2, 0.0, 1.0], [51, 255, 0, 255], 1)
    want = [round((0.2 + 51 / 255) / 2 * 255), round((0 + 1) / 2 * 255), round((1 + 0) / 2 * 255), 255]
    assert list(out) == want
    # currentSample 3: (c + 3 * acc/255) / 4
    out = orc.accumulate([1.0, 1.0, 1.0], [0, 0, 0, 255], 3)
    assert list(out) == [64, 64, 64, 255]


# ------------------------------------------------------------------ camera
def test_camera_model(orc):
    w = h = 256
    rays = orc.camera_rays([0, 3, 0], w, h, vfov=60.0).reshape(h, w, 6)
    # origin = camPos.zxy * (-1, 1, 1) = (-0, 0, 3)
    assert tuple(rays[0, 0, :3]) == (0.0, 0.0, 3.0)
    # the centre pixel looks straight down -z (no half-pixel offset)
    assert np.allclose(rays[128, 128, 3:], [0, 0, -1], atol=1e-7)
    # pixel (0, 0) is the top-left corner of a viewport 2*tan(30 deg) high
    vh = 2 * math.tan(math.radians(30))
    c = np.array([-vh / 2, vh / 2, -1.0])
    assert np.allclose(rays[0, 0, 3:], c / np.linalg.norm(c), atol=1e-6)
    # rows go down the image: y decreases with the row index
    assert rays[10, 128, 4] > rays[200, 128, 4]
