"""Cross-checks of the oracle's three traversals (CPU only): the shader's
visiting order with its line test (hit_bvh, ray-trace-compute.comp:263-311),
the ordered interval-culled traversal, and brute force over every triangle
(the triangle half of hit_scene, :222-235).  Any tree that encloses its
triangles gives the same closest hit, so the three must agree except where
two hits tie in t.  Also the MAX_STACK_DEPTH truncation facts."""
import numpy as np
import pytest

import oracle as O


def scene_of(rt, hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def near_tie(sc, ray, id_a, id_b):
    """True when the two triangles' t on this ray agree within 1e-6."""
    ts = []
    for i in (id_a, id_b):
        if i < 0:
            return False
        ok = O.hit_triangle(sc.triangles[i:i + 1], ray["o"], ray["d"], float(ray["tmin"]), float(ray["tmax"]))
        ts.append(ok[3])
    return abs(ts[0] - ts[1]) <= 1e-6 * max(1.0, abs(ts[0]))


def mixed_rays(rt, n, seed):
    a = rt.rays_shell(n // 2, seed, 3.0, 0.5)
    b = rt.rays_box(n - n // 2, seed + 1, 1.2)
    return np.concatenate([a, b])


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 3), (64, 4), (1024, 1)])
def test_three_traversals_agree(rt, n, seed):
    hs = rt.scene_soup(n, seed, threads=1)
    sc = scene_of(rt, hs)
    rays = mixed_rays(rt, 4096, 100 + n)
    ref = O.closest_hit(sc, rays, O.TRAV_REFERENCE)
    ordd = O.closest_hit(sc, rays, O.TRAV_ORDERED)
    brute = O.closest_hit(sc, rays, O.TRAV_BRUTE)
    hits = int((ref["id"] >= 0).sum())
    assert hits > (len(rays) // 50 if n >= 64 else 0)
    for other in (ordd, brute):
        diff = np.nonzero(ref["id"] != other["id"])[0]
        for i in diff:
            assert near_tie(sc, rays[i], int(ref["id"][i]), int(other["id"][i])), i
        same = ref["id"] == other["id"]
        for f in ("t", "u", "v"):
            assert np.array_equal(ref[f][same].view(np.uint32), other[f][same].view(np.uint32))


def test_ordered_traversal_visits_fewer_nodes(rt):
    hs = rt.scene_soup(1024, 1, threads=1)
    sc = scene_of(rt, hs)
    rays = mixed_rays(rt, 2048, 5)
    _, nv_ref, tt_ref, _ = O.closest_hit(sc, rays, O.TRAV_REFERENCE, stats=True)
    _, nv_ord, tt_ord, _ = O.closest_hit(sc, rays, O.TRAV_ORDERED, stats=True)
    assert tt_ord.sum() <= tt_ref.sum()
    # each ordered step reads one two-box record; the line test reads one box per node
    assert nv_ord.sum() < nv_ref.sum()


def test_icosphere_closed_surface(rt):
    # every ray from outside aimed at the interior crosses the closed surface
    hs = rt.scene_icosphere(3, 2)
    sc = scene_of(rt, hs)
    rays = rt.rays_shell(2000, 9, 3.0, 0.5)
    h = O.closest_hit(sc, rays, O.TRAV_ORDERED)
    assert (h["id"] >= 0).all()
    assert (h["t"] > 1.0).all() and (h["t"] < 3.0).all()
    assert ((h["u"] >= 0) & (h["v"] >= 0) & (h["u"] + h["v"] <= 1)).all()


def leaf_depths(nodes):
    depth = np.zeros(len(nodes), np.int64)
    for i in range(len(nodes)):          # children are numbered after parents
        if nodes["left"][i] >= 0:
            depth[nodes["left"][i]] = depth[i] + 1
            depth[nodes["right"][i]] = depth[i] + 1
    return depth


def rays_at_leaves(hs, leaves):
    """One ray per leaf, from outside the scene at that leaf's triangle centroid."""
    from oracle import RAY_DT
    rays = np.zeros(len(leaves), RAY_DT)
    for k, i in enumerate(leaves):
        t = hs.triangles[hs.nodes["object"][i]]
        c = (t["v0"].astype(np.float64) + t["v1"] + t["v2"]) / 3
        o = np.array([0.3, 5.0, 0.2])
        d = (c - o) / np.linalg.norm(c - o)
        rays["o"][k], rays["d"][k] = o, d
    rays["tmin"], rays["tmax"] = 0.001, 10000.0
    return rays


@pytest.mark.parametrize("n,truncates", [(8192, False), (8193, True)])
def test_max_stack_16_truncation_threshold(rt, n, truncates):
    """Median splits put the larger half on the right, which the shader's DFS
    expands first, so the leaf at the end of the all-right path at depth d is
    reached with d pending left siblings plus its own two pushes: d + 2
    entries.  Every leaf of the 8192-triangle tree sits at depth 13 (peak
    15); the 8193-triangle tree has its right-most leaf at depth 14 (peak
    16), which MAX_STACK_DEPTH 16 refuses (ray-trace-compute.comp:262,280)."""
    hs = rt.scene_soup(n, 21, threads=4)
    sc = scene_of(rt, hs)
    depth = leaf_depths(hs.nodes)
    leaves = np.nonzero(hs.nodes["object"] >= 0)[0]
    assert depth[leaves].max() == (14 if n == 8193 else 13)
    deepest = leaves[depth[leaves] == depth[leaves].max()]
    # the right child of the deepest split carries the larger index
    rays = rays_at_leaves(hs, np.concatenate([deepest[-1:], leaves[::257]]))
    _, _, _, tr = O.closest_hit(sc, rays, O.TRAV_REFERENCE, max_stack=16, stats=True, threads=8)
    assert bool(tr[0]) is truncates
    assert bool(tr.any()) is truncates
    _, _, _, tr0 = O.closest_hit(sc, rays, O.TRAV_REFERENCE, max_stack=0, stats=True, threads=8)
    assert not tr0.any()


def test_bug_compatible_sphere_rule(rt):
    """hit_scene resets closest_so_far before the sphere loop (RC:236): any
    sphere hit in range replaces a nearer triangle hit."""
    hs = rt.scene_soup(1024, 1, threads=1)
    sc = scene_of(rt, hs)
    rays = rt.rays_camera_tiles([0, 3, 0], 60.0, 64, 64, 8, 0, 4096)
    correct = O.closest_hit(sc, rays, O.TRAV_BRUTE, spheres=1)
    buggy = O.closest_hit(sc, rays, O.TRAV_BRUTE, spheres=2)
    ntri = len(hs.triangles)
    sph = buggy["id"] >= ntri
    # wherever the buggy rule reports a sphere, the correct rule has a hit no farther
    assert sph.any()
    assert (correct["t"][sph] <= buggy["t"][sph]).all()
    assert sph.sum() >= (correct["id"] >= ntri).sum()
    # the bug shows: a sphere wins although a triangle is nearer
    assert ((correct["id"] < ntri) & (correct["id"] >= 0) & sph).any()


def test_equal_sphere_roots_go_to_the_later_sphere(rt):
    """hit_sphere accepts root == tMax (ray-trace-compute.comp:192, `tMax <
    root` rejects), so in the sphere loop a later sphere at the same t
    replaces an earlier one."""
    hs = rt.scene_soup(16, 3)
    sp = np.zeros(3, O.SPHERE_DT)
    sp["s"] = [[5.0, 5.0, 5.0, 0.5], [5.0, 5.0, 5.0, 0.5], [9.0, 9.0, 9.0, 0.1]]
    sc = O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, sp)
    rays = np.zeros(1, O.RAY_DT)
    rays["o"], rays["d"], rays["tmin"], rays["tmax"] = [5.0, 5.0, 2.0], [0.0, 0.0, 1.0], 0.001, 10000.0
    for rule in (1, 2):
        h = O.closest_hit(sc, rays, O.TRAV_BRUTE, spheres=rule)
        assert h["id"][0] == len(hs.triangles) + 1 and h["t"][0] == np.float32(2.5)
