"""Pins the oracle's restatement of the GPU builder's linear BVH
(oracle/rt_oracle.c orc_lbvh; Karras, HPG 2012) by closed forms, brute
force over the augmented keys, exact box unions and traversal equivalence.
The reference has no such builder (its Bvh::createBvh, Bvh.h:141-209, is
pinned in test_bvh_builder.py)."""
import numpy as np
import pytest

import oracle as O


def tri_soup(n, seed, spread=1.0, size=0.05):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, O.TRI_DT)
    c = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    for k in ("v0", "v1", "v2"):
        t[k] = c + rng.uniform(-size, size, (n, 3)).astype(np.float32)
    return t


def tri_bounds(t):
    """Triangle boxes as the reference builder pads them (Bvh.h:96-97)."""
    v = np.stack([t["v0"], t["v1"], t["v2"]], axis=1)
    eps = np.float32(0.0001)
    return v.min(axis=1) - eps, v.max(axis=1) + eps


def leaf_ranges(nodes, n):
    """Sorted positions under every node (leaf k sits at node n-1+k)."""
    out = {}

    def walk(i):
        nd = nodes[i]
        if nd["object"] != -1:
            out[i] = [i - (n - 1)]
            return out[i]
        out[i] = walk(int(nd["left"])) + walk(int(nd["right"]))
        return out[i]

    import sys
    sys.setrecursionlimit(10000)
    walk(0)
    return out


def test_morton_closed_forms():
    assert O.morton(1, 0, 0) == 4 and O.morton(0, 1, 0) == 2 and O.morton(0, 0, 1) == 1
    for b in range(21):
        assert O.morton(1 << b, 0, 0) == 1 << (3 * b + 2)
        assert O.morton(0, 1 << b, 0) == 1 << (3 * b + 1)
        assert O.morton(0, 0, 1 << b) == 1 << (3 * b)
    m = (1 << 21) - 1
    assert O.morton(m, m, m) == (1 << 63) - 1
    assert O.morton(1 << 21, 0, 0) == 0            # only 21 bits per axis
    rng = np.random.default_rng(0)
    for _ in range(100):                            # interleaving is bitwise: OR-linear
        x1, x2 = (int(v) for v in rng.integers(0, 1 << 21, 2))
        y, z = (int(v) for v in rng.integers(0, 1 << 21, 2))
        assert O.morton(x1 | x2, y, z) == O.morton(x1, y, z) | O.morton(x2, 0, 0)


def test_single_triangle_is_one_leaf():
    t = tri_soup(1, 1)
    nodes, codes, order = O.lbvh(t)
    lo, hi = tri_bounds(t)
    assert len(nodes) == 1 and nodes[0]["object"] == 0 and nodes[0]["left"] == -1 and nodes[0]["right"] == -1
    assert np.array_equal(nodes["min"][0], lo[0]) and np.array_equal(nodes["max"][0], hi[0])
    assert codes[0] == 0                            # every axis has zero extent
    assert (nodes["min"][0] < t["v0"][0]).all() and (nodes["max"][0] > t["v0"][0]).all()   # strictly encloses


def test_pair_orders_by_code():
    t = tri_soup(2, 2)
    t["v0"][0] += 5.0; t["v1"][0] += 5.0; t["v2"][0] += 5.0        # triangle 0 far along +x, y, z
    nodes, codes, order = O.lbvh(t)
    assert list(order) == [1, 0]
    assert codes[0] == 0 and codes[1] == (1 << 63) - 1             # opposite corners of the bounds
    assert nodes[0]["left"] == 1 and nodes[0]["right"] == 2 and nodes[0]["object"] == -1
    assert nodes[1]["object"] == 1 and nodes[2]["object"] == 0


def test_eight_octants_give_a_complete_tree():
    t = np.zeros(8, O.TRI_DT)
    for i in range(8):
        c = np.array([(i >> 2) & 1, (i >> 1) & 1, i & 1], np.float32) * 2 - 1
        t["v0"][i] = c + [0.01, 0, 0]
        t["v1"][i] = c + [0, 0.01, 0]
        t["v2"][i] = c + [0, 0, 0.01]
    nodes, codes, order = O.lbvh(t)
    # corner i has cells (x,y,z) in {0, 2^21-1}: code = x-bits | y-bits | z-bits
    m = (1 << 21) - 1
    want = sorted(O.morton(m * ((i >> 2) & 1), m * ((i >> 1) & 1), m * (i & 1)) for i in range(8))
    assert list(codes) == want and list(order) == list(range(8))
    depth, todo = {0: 0}, [0]
    while todo:
        i = todo.pop()
        for ch in (nodes[i]["left"], nodes[i]["right"]):
            if ch != -1:
                depth[int(ch)] = depth[i] + 1
                todo.append(int(ch))
    assert all(depth[7 + k] == 3 for k in range(8))                 # every leaf at depth 3
    ranges = leaf_ranges(nodes, 8)
    assert sorted(ranges[int(nodes[0]["left"])]) == [0, 1, 2, 3]    # first split: x low / x high


@pytest.mark.parametrize("n,seed,degenerate", [(3, 1, False), (17, 2, False), (100, 3, False), (1000, 4, False),
                                               (37, 5, True), (300, 6, True)])
def test_structure_split_rule_and_boxes(n, seed, degenerate):
    t = tri_soup(n, seed)
    if degenerate:                                  # shared centres: runs of equal codes
        t["v0"][::2] = t["v0"][0]; t["v1"][::2] = t["v1"][0]; t["v2"][::2] = t["v2"][0]
    nodes, codes, order = O.lbvh(t)
    assert len(nodes) == 2 * n - 1
    assert sorted(order.tolist()) == list(range(n))
    keys = sorted(zip(codes.tolist(), order.tolist()))
    assert keys == list(zip(codes.tolist(), order.tolist()))       # (code, index) ascending
    ranges = leaf_ranges(nodes, n)
    assert len(ranges) == 2 * n - 1                                 # every node reached once
    lo, hi = tri_bounds(t)
    aug = [(int(codes[k]) << 32) | k for k in range(n)]
    for i, pos in ranges.items():
        f, l = min(pos), max(pos)
        assert sorted(pos) == list(range(f, l + 1))                  # contiguous sorted range
        tri = order[f:l + 1]
        assert np.array_equal(nodes["min"][i], lo[tri].min(axis=0))  # exact union of the leaves
        assert np.array_equal(nodes["max"][i], hi[tri].max(axis=0))
        if f == l:
            assert i == n - 1 + f and nodes[i]["object"] == order[f]
            continue
        assert i in (f, l) or i == 0                                # radix-tree numbering
        p = (aug[f] ^ aug[l]).bit_length() - 1
        left = ranges[int(nodes[i]["left"])]
        g = max(left)
        for m in range(f, l + 1):                                   # brute-force split rule
            assert ((aug[m] >> p) & 1 == 0) == (m <= g)


def test_lbvh_tree_traverses_like_brute_force(rt):
    hs = rt.scene_soup(1000, 8)
    nodes, _, _ = O.lbvh(hs.triangles)
    sc_l = O.Scene(hs.triangles, hs.materials, nodes, hs.lights)
    rays = np.concatenate([rt.rays_shell(1500, 9, 3.0, 0.5), rt.rays_box(1500, 10, 1.2)])
    a = O.closest_hit(sc_l, rays, O.TRAV_ORDERED)
    b = O.closest_hit(sc_l, rays, O.TRAV_BRUTE)
    same = a["id"] == b["id"]
    assert same.mean() > 0.999
    assert np.array_equal(a["t"][same].view(np.uint32), b["t"][same].view(np.uint32))


def test_codes_use_the_box_centre():
    """Box centre, not vertex mean: a triangle spanning [0, 0.5] per axis
    (its padded box [-0.0001, 0.5001] has centre 0.25 in binary32) sits at
    0.25 of the unit bounds, cell 2^19 on every axis (its vertex mean, 1/6,
    would give cell 349525)."""
    t = np.zeros(3, O.TRI_DT)
    t["v0"][1] = t["v1"][1] = t["v2"][1] = [1.0, 1.0, 1.0]      # triangle 1: the point (1,1,1)
    t["v1"][2] = [0.5, 0.5, 0.5]                                 # triangle 2: (0,0,0) (0.5,0.5,0.5) (0,0,0)
    nodes, codes, order = O.lbvh(t)
    assert list(order) == [0, 2, 1]
    assert list(codes) == [0, 7 << 57, (1 << 63) - 1]
    assert O.morton(1 << 19, 1 << 19, 1 << 19) == 7 << 57
