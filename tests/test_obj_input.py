"""OBJ input (rt_obj_load: GpuModel::getTriangles over Mesh::Mesh,
RtScene.h:13-30, mesh.cpp:96-139) and the reference's doge scene
(rtw_scene_doge: RtScene.h:44-101), checked against hand-written files and
against tests/golden/doge_triangles.npy, which
tests/golden/make_doge_fixture.py writes from the reference's models with
a reader of its own."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "doge_triangles.npy")
DOGE_DIR = "/root/reference/resources/models/doge_scene"


def write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def corners(t):
    return np.stack([t["v0"], t["v1"], t["v2"]], axis=1)


def test_faces_corner_forms_and_fans(rt, tmp_path):
    obj = ("# comment\r\no thing\r\nv 0 0 0\r\nv 1 0 0\r\nv 1 1 0\r\nv 0 1 0\r\n"
           "vt 0.5 0.5\r\nvn 0 0 1\r\ns off\r\nusemtl x\r\n"
           "f 1 2 3\r\nf 1/1 3/1 4/1\r\nf 1//1 2//1 3//1\r\nf 1/1/1 2/1/1 3/1/1 4/1/1\r\nf -4 -3 -2\r\n")
    t = rt.obj_load(write(tmp_path, obj), 7)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    want = [v[[0, 1, 2]], v[[0, 2, 3]], v[[0, 1, 2]], v[[0, 1, 2]], v[[0, 2, 3]], v[[0, 1, 2]]]
    assert np.array_equal(corners(t), np.array(want))
    assert (t["materialIndex"] == 7).all()


def test_numbers_round_like_strtod_then_float(rt, tmp_path):
    vals = ["-0.530557", "1e-3", "3.4028235e38", "0.1", "-0", "16777217"]
    obj = "v " + " ".join(vals[:3]) + "\nv " + " ".join(vals[3:]) + "\nv 0 0 0\nf 1 2 3\n"
    t = rt.obj_load(write(tmp_path, obj))
    want = np.array([float(x) for x in vals], np.float64).astype(np.float32)
    assert np.array_equal(t["v0"][0].view(np.uint32), want[:3].view(np.uint32))
    assert np.array_equal(t["v1"][0].view(np.uint32), want[3:].view(np.uint32))


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nf 1 2\n", "v 0 0 0\nf 1 2 3\n", "v 0 0 0\nf 0 1 1\n",
                                  "v 0 0\nf 1 1 1\n", "v 0 0 zz\n", "v 0 0 0\nf -2 1 1\n"])
def test_malformed_files_are_rejected(rt, tmp_path, text):
    with pytest.raises(rt.RtError) as e:
        rt.obj_load(write(tmp_path, text))
    assert e.value.status == rt.RT_ERR_INVALID and "line" in str(e.value)


def test_missing_file(rt, tmp_path):
    with pytest.raises(rt.RtError) as e:
        rt.obj_load(str(tmp_path / "absent.obj"))
    assert e.value.status == rt.RT_ERR_NOT_FOUND


def test_doge_fixture_shape():
    raw = np.load(GOLDEN)
    assert raw.shape == (1422, 48)
    mats = raw.view(np.uint32).reshape(-1, 12)[:, 11]
    runs = [(916, 0), (480, 0), (2, 1), (2, 2), (2, 0), (16, 0), (2, 0), (2, 3)]
    assert np.array_equal(mats, np.concatenate([np.full(n, m, np.uint32) for n, m in runs]))


@pytest.mark.skipif(not os.path.isdir(DOGE_DIR), reason="needs the reference's doge models")
def test_doge_scene_from_the_reference_models(rt):
    hs = rt.scene_doge(DOGE_DIR)
    raw = np.load(GOLDEN)
    assert np.array_equal(np.ascontiguousarray(hs.triangles).view(np.uint8).reshape(-1, 48)[:, :44], raw[:, :44])
    assert np.array_equal(hs.triangles["materialIndex"], raw.view(np.uint32).reshape(-1, 12)[:, 11])
    assert list(hs.materials["type"]) == [rt.RT_MAT_LAMBERTIAN] * 3 + [rt.RT_MAT_LIGHT, rt.RT_MAT_METAL, rt.RT_MAT_GLASS]
    assert np.allclose(hs.materials["albedo"], [[.3, .3, .3], [.9, .1, .1], [.1, .9, .1], [2, 2, 2], [1, 1, 1], [1, 1, 1]])
    assert list(hs.lights["triangleIndex"]) == [1420, 1421]               # the two light.obj triangles
    v0, v1 = hs.triangles["v0"][1420:], hs.triangles["v1"][1420:]
    area = np.linalg.norm(np.cross(v0.astype(np.float64), v1.astype(np.float64)), axis=1) * 0.5
    assert np.allclose(hs.lights["area"], area, rtol=1e-6)                # RtScene.h:92 (cross of v0 and v1)
    assert np.array_equal(hs.spheres["s"], np.array([[0.6, 1.0, -1.0, 0.6]], np.float32)) and hs.spheres["materialIndex"][0] == 5
    assert np.array_equal(hs.nodes.view(np.uint8), rt.bvh_build(hs.triangles, seed=1).view(np.uint8))
