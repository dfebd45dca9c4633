"""Writes tests/golden/doge_triangles.npy: the triangle records (48 bytes
each, the GpuModel::Triangle layout) of the reference's doge scene in the
order and with the material indices of src/ray-tracing/RtScene.h:66-88,
read from the reference's resources/models/doge_scene/*.obj by a small OBJ
reader of its own (positions and triangular faces, 1-based indices; the
files hold nothing else that matters).  Run in a container that has the
reference checked out:

    python tests/golden/make_doge_fixture.py /root/reference/resources/models/doge_scene
"""
import os
import sys

import numpy as np

PARTS = [("buff-doge.obj", 0), ("cheems.obj", 0), ("right.obj", 1), ("left.obj", 2), ("back.obj", 0),
         ("ceil.obj", 0), ("floor.obj", 0), ("light.obj", 3)]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "doge_triangles.npy")


def read_obj(path):
    pos, faces = [], []
    for raw in open(path, encoding="ascii"):
        tok = raw.split()
        if not tok:
            continue
        if tok[0] == "v":
            pos.append([float(t) for t in tok[1:4]])
        elif tok[0] == "f":
            idx = [int(t.split("/")[0]) for t in tok[1:]]
            assert len(idx) == 3, "the doge models hold triangles only"
            faces.append([i - 1 if i > 0 else len(pos) + i for i in idx])
    p = np.array(pos, np.float64).astype(np.float32)
    return p[np.array(faces, np.int64)]                 # [faces, 3 corners, xyz]


def main(models_dir):
    rec = []
    for name, mat in PARTS:
        corners = read_obj(os.path.join(models_dir, name))
        r = np.zeros((len(corners), 12), np.float32)
        for k in range(3):
            r[:, 4 * k:4 * k + 3] = corners[:, k]
        r[:, 11] = np.array([mat], np.uint32).view(np.float32)[0]
        rec.append(r)
    out = np.concatenate(rec).view(np.uint8).reshape(-1, 48)
    np.save(OUT, out)
    print(f"wrote {OUT}: {len(out)} triangles")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/resources/models/doge_scene")
