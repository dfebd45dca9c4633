"""Python face of the oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built from ``rt_oracle.c`` by ``oracle/Makefile``)
and exposes the CPU restatement of ``ray-trace-compute.comp`` on numpy record
arrays whose layouts equal the reference's std430 records
(``src/ray-tracing/GpuModels.h:26-63``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the library never does.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# ----------------------------------------------------------------- records
TRI_DT = np.dtype({"names": ["v0", "v1", "v2", "materialIndex"],
                   "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), "<u4"],
                   "offsets": [0, 16, 32, 44], "itemsize": 48})
MAT_DT = np.dtype({"names": ["type", "albedo"], "formats": ["<u4", ("<f4", 3)],
                   "offsets": [0, 16], "itemsize": 32})
NODE_DT = np.dtype({"names": ["min", "max", "left", "right", "object"],
                    "formats": [("<f4", 3), ("<f4", 3), "<i4", "<i4", "<i4"],
                    "offsets": [0, 16, 28, 32, 36], "itemsize": 48})
LIGHT_DT = np.dtype({"names": ["triangleIndex", "area"], "formats": ["<u4", "<f4"],
                     "offsets": [0, 4], "itemsize": 8})
SPHERE_DT = np.dtype({"names": ["s", "materialIndex"], "formats": [("<f4", 4), "<u4"],
                      "offsets": [0, 16], "itemsize": 32})
RAY_DT = np.dtype({"names": ["o", "tmin", "d", "tmax"],
                   "formats": [("<f4", 3), "<f4", ("<f4", 3), "<f4"],
                   "offsets": [0, 12, 16, 28], "itemsize": 32})
HIT_DT = np.dtype({"names": ["id", "t", "u", "v"], "formats": ["<i4", "<f4", "<f4", "<f4"],
                   "offsets": [0, 4, 8, 12], "itemsize": 16})

TRAV_REFERENCE, TRAV_ORDERED, TRAV_BRUTE = 0, 1, 2


class _Scene(ctypes.Structure):
    _fields_ = [("tris", ctypes.c_void_p), ("ntri", ctypes.c_uint32),
                ("mats", ctypes.c_void_p), ("nmat", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("nnode", ctypes.c_uint32),
                ("lights", ctypes.c_void_p), ("nlight", ctypes.c_uint32),
                ("spheres", ctypes.c_void_p), ("nsphere", ctypes.c_uint32)]


class _RenderParams(ctypes.Structure):
    _fields_ = [("vfov", ctypes.c_float), ("bounces", ctypes.c_uint32),
                ("max_stack", ctypes.c_uint32), ("spheres", ctypes.c_uint32),
                ("trav", ctypes.c_uint32)]


class _Ubo(ctypes.Structure):
    _fields_ = [("cam", ctypes.c_float * 3), ("time", ctypes.c_float),
                ("current_sample", ctypes.c_uint32), ("num_tri", ctypes.c_uint32),
                ("num_lights", ctypes.c_uint32), ("num_spheres", ctypes.c_uint32)]


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load(path: str):
    """Loads one build of the restatement and declares its entry points."""
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    u32 = ctypes.c_uint32
    L.orc_sizeof.restype = ctypes.c_int
    L.orc_step_rng.restype = u32
    L.orc_step_rng.argtypes = [u32]
    L.orc_rng_seed.restype = u32
    L.orc_rng_seed.argtypes = [u32, u32, u32]
    L.orc_rng_stream.argtypes = [u32, u32, u32, u32, P, P]
    L.orc_intersect_aabb.argtypes = [P, P, P, P, P]
    L.orc_tri_intersect.argtypes = [P, P, P, P, P]
    L.orc_hit_triangle.argtypes = [P, P, P, ctypes.c_float, ctypes.c_float, P]
    L.orc_hit_sphere.argtypes = [P, P, P, ctypes.c_float, ctypes.c_float, P]
    L.orc_closest_hit.argtypes = [P, P, u32, P, ctypes.c_int, u32, ctypes.c_int, P, P, P, ctypes.c_int]
    L.orc_scatter.argtypes = [P, P, P, P, ctypes.c_int, u32, P, P, P]
    L.orc_scatter.restype = ctypes.c_int
    L.orc_accumulate.argtypes = [P, P, u32, P]
    L.orc_render.argtypes = [P, P, P, u32, u32, u32, P, P, P, ctypes.c_int]
    L.orc_render.restype = ctypes.c_uint64
    L.orc_camera_rays.argtypes = [P, ctypes.c_float, u32, u32, P]
    L.orc_morton.restype = ctypes.c_uint64
    L.orc_morton.argtypes = [u32, u32, u32]
    L.orc_lbvh.restype = ctypes.c_int
    L.orc_lbvh.argtypes = [P, u32, P, P, P]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using_library(path: str):
    """Runs this module's functions on another build of the restatement (a
    mutant compiled by a test, to show that a known-answer test detects a
    planted mistake), then restores the committed one."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def _ptr(a: np.ndarray | None):
    if a is None or len(a) == 0:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Scene:
    """Host arrays of one scene in the reference's record layouts."""

    def __init__(self, triangles, materials, nodes, lights=None, spheres=None):
        self.triangles = np.ascontiguousarray(triangles, dtype=TRI_DT)
        self.materials = np.ascontiguousarray(materials, dtype=MAT_DT)
        self.nodes = np.ascontiguousarray(nodes, dtype=NODE_DT)
        self.lights = np.ascontiguousarray(lights if lights is not None else np.zeros(0, LIGHT_DT), dtype=LIGHT_DT)
        self.spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE_DT), dtype=SPHERE_DT)

    def _c(self) -> _Scene:
        return _Scene(_ptr(self.triangles), len(self.triangles), _ptr(self.materials), len(self.materials),
                      _ptr(self.nodes), len(self.nodes), _ptr(self.lights), len(self.lights),
                      _ptr(self.spheres), len(self.spheres))


# --------------------------------------------------------------------- RNG
def rng_stream(x: int, y: int, sample: int, n: int):
    """Draws 0..n-1 of random() for invocation (x, y) at currentSample."""
    out = np.zeros(n, np.float32)
    words = np.zeros(n, np.uint32)
    lib().orc_rng_stream(x, y, sample, n, out.ctypes.data, words.ctypes.data)
    return out, words


def rng_seed(x: int, y: int, sample: int) -> int:
    return lib().orc_rng_seed(x, y, sample)


# ----------------------------------------------------------- intersections
def intersect_aabb(ro, rd, bmin, bmax):
    a = [np.ascontiguousarray(v, np.float32) for v in (ro, rd, bmin, bmax)]
    out = np.zeros(2, np.float32)
    lib().orc_intersect_aabb(*(v.ctypes.data for v in a), out.ctypes.data)
    return out


def tri_intersect(ro, rd, tri):
    ro = np.ascontiguousarray(ro, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    tri = np.ascontiguousarray(np.atleast_1d(tri), TRI_DT)
    tuv = np.zeros(3, np.float32)
    n = np.zeros(3, np.float32)
    lib().orc_tri_intersect(ro.ctypes.data, rd.ctypes.data, tri.ctypes.data, tuv.ctypes.data, n.ctypes.data)
    return tuv, n


def hit_triangle(tri, ro, rd, tmin=0.001, tmax=10000.0):
    """Returns (hit, p[3], normal[3], t, backFaceInt, materialIndex)."""
    ro = np.ascontiguousarray(ro, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    tri = np.ascontiguousarray(np.atleast_1d(tri), TRI_DT)
    out = np.zeros(9, np.float32)
    ok = lib().orc_hit_triangle(tri.ctypes.data, ro.ctypes.data, rd.ctypes.data, tmin, tmax, out.ctypes.data)
    return bool(ok), out[0:3].copy(), out[3:6].copy(), float(out[6]), int(out[7]), int(out[8])


def hit_sphere(sph, ro, rd, tmin=0.001, tmax=10000.0):
    ro = np.ascontiguousarray(ro, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    sph = np.ascontiguousarray(np.atleast_1d(sph), SPHERE_DT)
    out = np.zeros(9, np.float32)
    ok = lib().orc_hit_sphere(sph.ctypes.data, ro.ctypes.data, rd.ctypes.data, tmin, tmax, out.ctypes.data)
    return bool(ok), out[0:3].copy(), out[3:6].copy(), float(out[6]), int(out[7]), int(out[8])


def closest_hit(scene: Scene, rays: np.ndarray, traversal: int = TRAV_REFERENCE, max_stack: int = 0,
                spheres: int = 0, stats: bool = False, threads: int = 1):
    rays = np.ascontiguousarray(rays, RAY_DT)
    n = len(rays)
    hits = np.zeros(n, HIT_DT)
    nv = np.zeros(n, np.uint32) if stats else None
    tt = np.zeros(n, np.uint32) if stats else None
    tr = np.zeros(n, np.uint8) if stats else None
    sc = scene._c()
    rc = lib().orc_closest_hit(ctypes.byref(sc), _ptr(rays), n, _ptr(hits), traversal, max_stack, spheres,
                               _ptr(nv), _ptr(tt), _ptr(tr), threads)
    if rc != 0:
        raise ValueError("oracle rejected the scene")
    if stats:
        return hits, nv, tt, tr
    return hits


def scatter(scene: Scene, rin_dir, p, normal, back, mat, state):
    """Runs scatter() once; returns (emits, albedo[3], dir[3], new_state)."""
    a = [np.ascontiguousarray(v, np.float32) for v in (rin_dir, p, normal)]
    st = ctypes.c_uint32(state)
    alb = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    sc = scene._c()
    e = lib().orc_scatter(ctypes.byref(sc), a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, int(back),
                          int(mat), ctypes.byref(st), alb.ctypes.data, d.ctypes.data)
    return bool(e), alb, d, st.value


def accumulate(color3, accum4, current_sample):
    c = np.ascontiguousarray(color3, np.float32)
    a = np.ascontiguousarray(accum4, np.uint8)
    out = np.zeros(4, np.uint8)
    lib().orc_accumulate(c.ctypes.data, a.ctypes.data, current_sample, out.ctypes.data)
    return out


def make_ubo(cam_pos, current_sample=0, scene: Scene | None = None, time=0.0) -> _Ubo:
    u = _Ubo()
    u.cam[0], u.cam[1], u.cam[2] = [float(v) for v in cam_pos]
    u.time = time
    u.current_sample = current_sample
    if scene is not None:
        u.num_tri, u.num_lights, u.num_spheres = len(scene.triangles), len(scene.lights), len(scene.spheres)
    return u


def render(scene: Scene, cam_pos, w, h, current_sample=0, vfov=30.0, bounces=2, max_stack=0, spheres=0,
           traversal=TRAV_REFERENCE, accum=None, row_step=1, threads=1, colors=False):
    """One frame of main(); returns (target RGBA8 [h,w,4], rays, colours?)."""
    ubo = make_ubo(cam_pos, current_sample, scene)
    prm = _RenderParams(vfov, bounces, max_stack, spheres, traversal)
    target = np.zeros((h, w, 4), np.uint8)
    acc = np.zeros((h, w, 4), np.uint8) if accum is None else np.ascontiguousarray(accum, np.uint8)
    col = np.zeros((h, w, 3), np.float32) if colors else None
    sc = scene._c()
    rays = lib().orc_render(ctypes.byref(sc), ctypes.byref(ubo), ctypes.byref(prm), w, h, row_step,
                            target.ctypes.data, acc.ctypes.data, col.ctypes.data if colors else None, threads)
    if colors:
        return target, int(rays), col
    return target, int(rays)


def render_frames(scene: Scene, cam_pos, w, h, frames, first_sample=0, **kw):
    """`frames` frames with the target->accumulation copy between them."""
    accum = np.zeros((h, w, 4), np.uint8)
    total = 0
    target = accum
    for k in range(frames):
        target, rays = render(scene, cam_pos, w, h, current_sample=first_sample + k, accum=accum, **kw)
        accum = target.copy()
        total += rays
    return target, total


def camera_rays(cam_pos, w, h, vfov=30.0, current_sample=0):
    ubo = make_ubo(cam_pos, current_sample)
    out = np.zeros((h * w, 6), np.float32)
    lib().orc_camera_rays(ctypes.byref(ubo), vfov, w, h, out.ctypes.data)
    return out


def morton(x: int, y: int, z: int) -> int:
    """63-bit Morton code of three 21-bit cells (x in bits 3b+2)."""
    return int(lib().orc_morton(x, y, z))


def lbvh(triangles: np.ndarray):
    """The linear BVH of the GPU builder: (nodes [2n-1], sorted codes [n], order [n])."""
    tris = np.ascontiguousarray(triangles, dtype=TRI_DT)
    n = len(tris)
    if n == 0:
        raise ValueError("lbvh: no triangles")
    nodes = np.zeros(2 * n - 1, NODE_DT)
    codes = np.zeros(n, np.uint64)
    order = np.zeros(n, np.uint32)
    if lib().orc_lbvh(_ptr(tris), n, _ptr(codes), _ptr(order), _ptr(nodes)) != 0:
        raise MemoryError("lbvh: allocation failed")
    return nodes, codes, order
