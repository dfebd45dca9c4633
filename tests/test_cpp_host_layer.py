"""The C++ host layer (include/rtamd.hpp) and the host program built on it
(vulkan-compute-ray-tracing_amd/rtamd_frame_loop, the reference's main.cpp
frame loop without a window).

CPU: the header compiles with g++ -pedantic -Werror, its host-side helpers
give the library's own tree and lights byte for byte, failing calls throw
rtamd::Error, and without a GPU the program fails loudly (exit 1, the
rt_create error), never falling back to the CPU.
GPU: the program's target image is byte-identical to the Python binding's
frame loop on the same scene, camera and frame count, and its PNG is the
display pass of that target."""
import os
import shutil
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vulkan-compute-ray-tracing_amd")
BIN = os.path.join(PKG, "rtamd_frame_loop")
CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "hpp_host_check.cpp")


def _gpu_present() -> bool:
    import torch
    return torch.cuda.device_count() > 0      # counts devices without initialising one


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        chunks.append((data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]))
        pos += 12 + n
    w, h = int.from_bytes(chunks[0][1][0:4], "big"), int.from_bytes(chunks[0][1][4:8], "big")
    raw = zlib.decompress(b"".join(body for typ, body in chunks if typ == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * 4)
    assert (rows[:, 0] == 0).all()          # the writer uses filter 0 on every row
    return rows[:, 1:].reshape(h, w, 4)


@pytest.fixture(scope="module")
def hpp_check(rt, tmp_path_factory):
    rt.lib()                                  # librtamd.so is built
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not present")
    out = str(tmp_path_factory.mktemp("hpp") / "hpp_check")
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           CHECK_SRC, "-o", out, "-L", PKG, "-lrtamd", f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_header_layer_host_checks(hpp_check):
    args = [hpp_check] + ([] if _gpu_present() else ["--no-device"])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "hpp ok" in r.stdout


def test_frame_loop_program_is_built(rt):
    rt.lib()
    assert os.access(BIN, os.X_OK), "rtamd_frame_loop is built by build.py next to librtamd.so"


def test_frame_loop_program_rejects_bad_arguments(rt):
    for args in ([], ["16", "16"], ["16", "16", "0", "x.png"], ["16", "x", "1", "x.png"],
                 ["16", "16", "1", "x.png", "--cam", "1", "2"], ["16", "16", "1", "x.png", "--what"]):
        r = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage:" in r.stderr, (args, r.stderr)


def test_frame_loop_program_fails_loudly_without_a_gpu(rt, tmp_path):
    if _gpu_present():
        pytest.skip("a GPU is present")
    out = str(tmp_path / "x.png")
    r = subprocess.run([BIN, "16", "16", "1", out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "rt_create" in r.stderr, r.stderr
    assert not os.path.exists(out)


@pytest.mark.gpu
def test_frame_loop_program_matches_the_binding(rt, request, tmp_path):
    w, h, frames, cam = 96, 80, 3, [0.0, 3.0, 0.0]
    png, raw = str(tmp_path / "frame.png"), str(tmp_path / "target.rgba")
    print("case: rtamd_frame_loop soup 1024 seed 1, 96x80, 3 frames", flush=True)
    r = subprocess.run([BIN, str(w), str(h), str(frames), png, "--soup", "1024", "1", "--cam", *map(str, cam),
                        "--vfov", "60", "--bounces", "4", "--raw", raw], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "currentSample 3" in r.stdout

    gpu_ctx = request.getfixturevalue("gpu_ctx")   # this process touches the GPU after the child has run
    scene = gpu_ctx.upload(rt.scene_soup(1024, 1))
    loop = rt.FrameLoop(gpu_ctx, scene, w, h, vfov=60.0, bounces=4)
    try:
        for _ in range(frames):
            loop.step(cam)
        target = loop.image("target")
    finally:
        loop.close()
        scene.release()
    got = np.fromfile(raw, np.uint8).reshape(h, w, 4)
    assert np.array_equal(got, target)
    disp = _read_png(png)
    lut = rt.gamma_table()
    assert np.array_equal(disp[..., :3], lut[target[..., :3]]) and (disp[..., 3] == 255).all()
