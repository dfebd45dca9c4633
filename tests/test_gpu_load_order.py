"""The binding works whichever of rtamd and torch touches the GPU first:
a fresh process creates an rtamd context before importing torch, then
uses torch tensors on the device with the library."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path[:0] = [sys.argv[1]]
import rtamd                              # torch not imported yet
ctx = rtamd.Context(0)
hs = rtamd.scene_soup(256, 3)
scene = ctx.upload(hs)
import torch
rays = rtamd.rays_to_device(rtamd.rays_box(4096, 5, 1.0))
hits = rtamd.hits_from_device(ctx.closest_hit(scene, rays))
assert (hits["id"] >= 0).any()
print("load order ok", int((hits["id"] >= 0).sum()))
"""


def test_context_before_torch():
    print("case: rtamd context created before torch is imported (child process)", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "vulkan-compute-ray-tracing_amd")],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "load order ok" in r.stdout
