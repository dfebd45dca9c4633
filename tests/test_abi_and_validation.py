"""C-ABI surface (no GPU needed): every symbol include/*.h declares is
exported by librtamd.so, and scene validation behaves as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:rt|rtw)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(rt):
    L = rt.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(rt.EXPORTS)


def test_status_strings(rt):
    L = rt.lib()
    assert L.rt_status_string(0) == b"ok"
    assert L.rt_status_string(1) == b"invalid argument"


def test_record_layouts_match_reference(rt):
    # GpuModels.h:26-63 std430 sizes and offsets
    assert rt.TRI_DT.itemsize == 48 and rt.TRI_DT.fields["materialIndex"][1] == 44
    assert rt.MAT_DT.itemsize == 32 and rt.MAT_DT.fields["albedo"][1] == 16
    assert rt.NODE_DT.itemsize == 48
    assert [rt.NODE_DT.fields[k][1] for k in ("min", "max", "left", "right", "object")] == [0, 16, 28, 32, 36]
    assert rt.LIGHT_DT.itemsize == 8 and rt.SPHERE_DT.itemsize == 32
    assert ctypes.sizeof(rt.Ubo) == 32 and rt.Ubo.currentSample.offset == 16


def test_parameter_structs_match_the_header(rt):
    """ctypes mirrors of the C-ABI's parameter structs (include/rtamd.h)."""
    import ctypes
    sizes = {"Tuning": 32, "QueryParams": 16, "RenderParams": 32, "RenderStats": 40, "BatchItem": 32,
             "BatchConfig": 32, "ServerConfig": 40, "SceneInfo": 48, "QueryStats": 24}
    for name, size in sizes.items():
        assert ctypes.sizeof(getattr(rt, name)) == size, name
    assert rt.Tuning.tree.offset == 16 and rt.Tuning.steal.offset == 24 and rt.RenderStats.total_ms.offset == 24


def test_oracle_layouts_agree(rt, orc):
    for a, b in [(rt.TRI_DT, orc.TRI_DT), (rt.MAT_DT, orc.MAT_DT), (rt.NODE_DT, orc.NODE_DT),
                 (rt.LIGHT_DT, orc.LIGHT_DT), (rt.SPHERE_DT, orc.SPHERE_DT), (rt.RAY_DT, orc.RAY_DT)]:
        assert a == b
    for k, size in enumerate([48, 32, 48, 8, 32, 32, 16]):
        assert orc.lib().orc_sizeof(k) == size


def small_scene(rt):
    return rt.scene_soup(64, 9, threads=1)


def test_valid_scene_takes_fast_path(rt):
    rc, info, msg = rt.scene_check(small_scene(rt))
    assert rc == 0 and info.fast_path == 1 and info.enclosing == 1
    assert info.reachable_leaves == 64 and info.tree_height == 6


def _set(attr, i, field, value):
    def f(s):
        getattr(s, attr)[field][i] = value
    return f


@pytest.mark.parametrize("mutate,needle", [
    (lambda s: setattr(s, "triangles", s.triangles[:0]), "no triangles"),
    (lambda s: setattr(s, "nodes", s.nodes[:0]), "no BVH nodes"),
    (_set("triangles", 3, "materialIndex", 99), "material out of range"),
    (_set("nodes", 0, "left", 10_000), "child index out of range"),
    (_set("nodes", 0, "left", -2), "child index out of range"),
    (_set("nodes", 5, "object", 64), "object index out of range"),
    (_set("materials", 1, "type", 7), "type > 3"),
    (_set("lights", 0, "triangleIndex", 1000), "light 0"),
    (_set("spheres", 1, "materialIndex", 4), "sphere 1"),
])
def test_invalid_scenes_are_rejected(rt, mutate, needle):
    s = small_scene(rt)
    mutate(s)
    rc, _, msg = rt.scene_check(s)
    assert rc == rt.RT_ERR_INVALID and needle in msg


def test_cycle_and_shared_subtree_rejected(rt):
    s = small_scene(rt)
    inner = np.nonzero(s.nodes["object"] < 0)[0]
    s.nodes["left"][inner[3]] = 0                       # back edge to the root
    rc, _, msg = rt.scene_check(s)
    assert rc == rt.RT_ERR_INVALID and "not a tree" in msg


def test_non_enclosing_tree_uses_reference_traversal(rt):
    s = small_scene(rt)
    leaf = np.nonzero(s.nodes["object"] >= 0)[0][0]
    s.nodes["max"][leaf] = s.nodes["min"][leaf]          # box no longer holds its triangle
    rc, info, _ = rt.scene_check(s)
    assert rc == 0 and info.enclosing == 0 and info.fast_path == 0


def test_node_with_triangle_and_children_is_valid_but_not_packed(rt):
    s = small_scene(rt)
    s.nodes["object"][0] = 0
    rc, info, _ = rt.scene_check(s)
    assert rc == 0 and info.fast_path == 0
