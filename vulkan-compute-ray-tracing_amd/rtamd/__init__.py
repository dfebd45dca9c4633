"""ctypes binding of librtamd.so (include/rtamd.h, include/rtamd_workloads.h).

The library is the product; this module only marshals arguments.  Device
buffers are torch tensors on a ROCm device (torch supplies HBM and streams);
host records are numpy arrays with the reference's std430 layouts
(src/ray-tracing/GpuModels.h:26-63).  There is no CPU fallback: if
librtamd.so is missing or the GPU is absent the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the in-tree build; RTAMD_LIB names another build of the same library
LIB_PATH = os.environ.get("RTAMD_LIB") or os.path.join(PKG_DIR, "librtamd.so")

RT_OK, RT_ERR_INVALID, RT_ERR_HIP, RT_ERR_NOMEM, RT_ERR_UNSUPPORTED, RT_ERR_BUSY, RT_ERR_CANCELLED, RT_ERR_NOT_FOUND = range(8)
RT_MAT_LIGHT, RT_MAT_LAMBERTIAN, RT_MAT_METAL, RT_MAT_GLASS = range(4)
SPHERES_OFF, SPHERES_NEAREST, SPHERES_SCENE = range(3)
TREE_GIVEN, TREE_DEVICE = range(2)
ORDER_AUTO, ORDER_GIVEN, ORDER_SORTED = range(3)
STEAL_NEVER = 65
TRAVERSE_AUTO, TRAVERSE_PACKED, TRAVERSE_REFERENCE = range(3)
LAYOUT_AUTO, LAYOUT_F32, LAYOUT_Q16, LAYOUT_W4 = range(4)

# --------------------------------------------------------------- records
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


class SceneDesc(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_void_p), ("num_triangles", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("num_materials", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("num_nodes", ctypes.c_uint32),
                ("lights", ctypes.c_void_p), ("num_lights", ctypes.c_uint32),
                ("spheres", ctypes.c_void_p), ("num_spheres", ctypes.c_uint32)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_uint32), ("num_nodes", ctypes.c_uint32),
                ("num_lights", ctypes.c_uint32), ("num_spheres", ctypes.c_uint32),
                ("num_materials", ctypes.c_uint32), ("tree_height", ctypes.c_uint32),
                ("reachable_leaves", ctypes.c_uint32), ("enclosing", ctypes.c_uint32),
                ("fast_path", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64)]


class QueryParams(ctypes.Structure):
    _fields_ = [("traversal", ctypes.c_uint32), ("max_stack", ctypes.c_uint32),
                ("spheres", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("vfov_deg", ctypes.c_float), ("bounces", ctypes.c_uint32),
                ("max_stack", ctypes.c_uint32), ("spheres", ctypes.c_uint32),
                ("traversal", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class Tuning(ctypes.Structure):
    _fields_ = [("node_layout", ctypes.c_uint32), ("refill", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("stack_lds", ctypes.c_uint32),
                ("tree", ctypes.c_uint32), ("ray_order", ctypes.c_uint32), ("steal", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class Ubo(ctypes.Structure):
    _fields_ = [("camPos", ctypes.c_float * 3), ("time", ctypes.c_float),
                ("currentSample", ctypes.c_uint32), ("numTriangles", ctypes.c_uint32),
                ("numLights", ctypes.c_uint32), ("numSpheres", ctypes.c_uint32)]


class BatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("rays_dev", ctypes.c_void_p), ("hits_dev", ctypes.c_void_p),
                ("count", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("stack_lds", ctypes.c_uint32), ("regroup", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("regroup_interval", ctypes.c_uint32), ("node_layout", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def batch_config(block=256, stack_lds=0, regroup=0, refill=32, blocks_per_cu=0, use_graph=1, regroup_interval=0,
                 node_layout=3) -> BatchConfig:
    return BatchConfig(block, stack_lds, regroup, refill, blocks_per_cu, use_graph, regroup_interval, node_layout)


class RenderStats(ctypes.Structure):
    _fields_ = [("primary_rays", ctypes.c_uint64), ("secondary_rays", ctypes.c_uint64),
                ("bounces", ctypes.c_uint32), ("timed", ctypes.c_uint32)] + \
               [(k, ctypes.c_float) for k in ("total_ms", "trace_primary_ms", "trace_secondary_ms", "shade_ms")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class QueryStats(ctypes.Structure):
    _fields_ = [("timed", ctypes.c_uint32), ("sorted", ctypes.c_uint32)] + \
               [(k, ctypes.c_float) for k in ("order_ms", "trace_ms", "spheres_ms", "total_ms")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("slots", ctypes.c_uint32), ("max_batch_rays", ctypes.c_uint32),
                ("max_render_pixels", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("deadline_ms", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "launches", "scenes_added",
                                               "scenes_evicted", "scenes_refused", "rays", "resident_bytes")] + \
               [(k, ctypes.c_double) for k in ("p50_ms", "p99_ms", "mean_ms", "max_ms")] + \
               [(k, ctypes.c_uint64) for k in ("peak_resident_bytes", "cancelled")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeParams(ctypes.Structure):
    _fields_ = [("rate_per_s", ctypes.c_double), ("duration_s", ctypes.c_double),
                ("render_fraction", ctypes.c_double), ("swap_interval_s", ctypes.c_double),
                ("query_min", ctypes.c_uint32), ("query_max", ctypes.c_uint32), ("render_w", ctypes.c_uint32),
                ("render_h", ctypes.c_uint32), ("render_bounces", ctypes.c_uint32), ("swap_triangles", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("threads", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    _fields_ = [("stats", ServerStats), ("wall_s", ctypes.c_double), ("requests", ctypes.c_uint64),
                ("queries", ctypes.c_uint64), ("renders", ctypes.c_uint64), ("swaps", ctypes.c_uint64)]


class ServeSample(ctypes.Structure):
    _fields_ = [("hits", ctypes.c_void_p), ("image", ctypes.c_void_p), ("scene_id", ctypes.c_uint64),
                ("arrival", ctypes.c_uint64), ("kind", ctypes.c_uint32), ("count", ctypes.c_uint32)]


class ServeSampling(ctypes.Structure):
    _fields_ = [("every", ctypes.c_uint32), ("capacity", ctypes.c_uint32), ("samples", ctypes.POINTER(ServeSample)),
                ("taken", ctypes.c_uint32), ("swaps_capacity", ctypes.c_uint32),
                ("swap_ids", ctypes.POINTER(ctypes.c_uint64)), ("swaps_taken", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


SERVE_BATCHED, SERVE_FIFO = 0, 1


class WScene(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_void_p), ("num_triangles", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("num_materials", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("num_nodes", ctypes.c_uint32),
                ("lights", ctypes.c_void_p), ("num_lights", ctypes.c_uint32),
                ("spheres", ctypes.c_void_p), ("num_spheres", ctypes.c_uint32)]


class RtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


_lib = None

# Every symbol include/rtamd.h and include/rtamd_workloads.h declare.
EXPORTS = [
    "rt_create", "rt_destroy", "rt_last_error", "rt_status_string", "rt_scene_upload", "rt_scene_info_get",
    "rt_scene_check", "rt_tuning_set", "rt_tuning_get",
    "rt_scene_release", "rt_closest_hit", "rt_closest_hit_stats", "rt_closest_hit_noalloc", "rt_render",
    "rt_render_frames",
    "rt_render_last_rays", "rt_render_set_timing", "rt_render_last_stats", "rt_render_set_counting",
    "rt_render_last_counts", "rt_query_last_stats", "rt_bvh_build", "rt_bvh_build_seeded", "rt_bvh_build_device", "rt_scene_lights",
    "rt_frame_loop_create", "rt_frame_loop_destroy", "rt_frame_loop_step", "rt_frame_loop_display",
    "rt_frame_loop_images", "rt_frame_loop_read", "rt_postprocess_gamma", "rt_gamma_table", "rt_image_write",
    "rt_batch_create", "rt_batch_set_config", "rt_batch_get_config", "rt_batch_launch", "rt_batch_autotune",
    "rt_batch_destroy",
    "rt_host_alloc", "rt_host_free", "rt_server_create", "rt_server_destroy", "rt_server_scene_add", "rt_server_scene_remove", "rt_server_scene_list",
    "rt_server_submit_query", "rt_server_submit_render", "rt_server_wait", "rt_server_cancel", "rt_server_drain", "rt_server_stats_get",
    "rt_server_stats_reset", "rt_server_last_error", "rtw_serve_poisson", "rtw_serve_poisson_sampled",
    "rtw_scene_soup", "rtw_scene_icosphere", "rtw_scene_clusters", "rtw_scene_doge", "rtw_scene_free",
    "rt_obj_load", "rt_obj_last_error", "rt_free", "rtw_rays_camera_tiles",
    "rtw_rays_shell", "rtw_rays_box", "rtw_log_uniform",
]


def lib():
    """Loads librtamd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run vulkan-compute-ray-tracing_amd/build.py")
    # torch ships its own HIP runtime.  With librtamd.so loaded and a context
    # created first, torch's later CUDA initialisation found no GPU; with
    # torch loaded first both share one runtime.  So torch, when installed,
    # is always imported before the library is loaded.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64
    sig = {
        "rt_create": (i32, [ctypes.POINTER(P), i32]),
        "rt_destroy": (i32, [P]),
        "rt_last_error": (ctypes.c_char_p, [P]),
        "rt_status_string": (ctypes.c_char_p, [i32]),
        "rt_scene_upload": (i32, [P, ctypes.POINTER(SceneDesc), ctypes.POINTER(P)]),
        "rt_scene_info_get": (i32, [P, ctypes.POINTER(SceneInfo)]),
        "rt_scene_release": (i32, [P, P]),
        "rt_scene_check": (i32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(SceneInfo), ctypes.c_char_p, ctypes.c_size_t]),
        "rt_tuning_set": (i32, [P, ctypes.POINTER(Tuning)]),
        "rt_tuning_get": (i32, [P, ctypes.POINTER(Tuning)]),
        "rt_closest_hit": (i32, [P, P, P, u32, P, ctypes.POINTER(QueryParams), P]),
        "rt_closest_hit_stats": (i32, [P, P, P, u32, P, P, P, P]),
        "rt_closest_hit_noalloc": (i32, [P, P, P, u32, P, ctypes.POINTER(QueryParams), P]),
        "rt_render": (i32, [P, P, ctypes.POINTER(Ubo), ctypes.POINTER(RenderParams), P, P, u32, u32, P]),
        "rt_render_frames": (i32, [P, P, ctypes.POINTER(Ubo), ctypes.POINTER(RenderParams), u32, P, P, u32, u32, P]),
        "rt_render_last_rays": (i32, [P, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "rt_render_set_timing": (i32, [P, i32]),
        "rt_render_last_stats": (i32, [P, ctypes.POINTER(RenderStats)]),
        "rt_render_set_counting": (i32, [P, i32]),
        "rt_render_last_counts": (i32, [P, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "rt_query_last_stats": (i32, [P, ctypes.POINTER(QueryStats)]),
        "rt_batch_create": (i32, [P, ctypes.POINTER(BatchItem), u32, ctypes.POINTER(BatchConfig), ctypes.POINTER(P)]),
        "rt_batch_set_config": (i32, [P, P, ctypes.POINTER(BatchConfig)]),
        "rt_batch_get_config": (i32, [P, ctypes.POINTER(BatchConfig)]),
        "rt_batch_launch": (i32, [P, P, P]),
        "rt_batch_autotune": (i32, [P, P, u32, ctypes.c_char_p, ctypes.POINTER(BatchConfig)]),
        "rt_batch_destroy": (i32, [P, P]),
        "rt_host_alloc": (i32, [ctypes.c_size_t, ctypes.POINTER(P)]),
        "rt_host_free": (i32, [P]),
        "rt_server_create": (i32, [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
        "rt_server_destroy": (i32, [P]),
        "rt_server_scene_add": (i32, [P, ctypes.POINTER(SceneDesc), ctypes.POINTER(u64)]),
        "rt_server_scene_remove": (i32, [P, u64]),
        "rt_server_scene_list": (i32, [P, P, u32, ctypes.POINTER(u32)]),
        "rt_server_submit_query": (i32, [P, u64, P, u32, P, P]),
        "rt_server_submit_render": (i32, [P, u64, ctypes.POINTER(Ubo), ctypes.POINTER(RenderParams), u32, u32, P, P]),
        "rt_server_wait": (i32, [P, u64, i32, ctypes.POINTER(i32)]),
        "rt_server_cancel": (i32, [P, u64]),
        "rt_server_drain": (i32, [P]),
        "rt_server_stats_get": (i32, [P, ctypes.POINTER(ServerStats)]),
        "rt_server_stats_reset": (i32, [P]),
        "rt_server_last_error": (ctypes.c_char_p, [P]),
        "rtw_serve_poisson": (i32, [P, ctypes.POINTER(ServeParams), ctypes.POINTER(ServeResult)]),
        "rtw_serve_poisson_sampled": (i32, [P, ctypes.POINTER(ServeParams), ctypes.POINTER(ServeSampling),
                                            ctypes.POINTER(ServeResult)]),
        "rt_bvh_build": (i32, [P, u32, P]),
        "rt_bvh_build_seeded": (i32, [P, u32, u32, u32, P]),
        "rt_bvh_build_device": (i32, [P, P, u32, P, P]),
        "rt_frame_loop_create": (i32, [P, P, ctypes.POINTER(RenderParams), u32, u32, ctypes.POINTER(P)]),
        "rt_frame_loop_destroy": (i32, [P]),
        "rt_frame_loop_step": (i32, [P, ctypes.POINTER(ctypes.c_float), ctypes.c_float, i32, P]),
        "rt_frame_loop_display": (i32, [P, P]),
        "rt_frame_loop_images": (i32, [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(u32)]),
        "rt_frame_loop_read": (i32, [P, i32, P, P]),
        "rt_postprocess_gamma": (i32, [P, P, P, u32, u32, P]),
        "rt_gamma_table": (i32, [P]),
        "rt_image_write": (i32, [ctypes.c_char_p, P, u32, u32]),
        "rt_scene_lights": (i32, [P, u32, P, u32, P, u32, ctypes.POINTER(u32)]),
        "rtw_scene_soup": (i32, [u32, u64, u32, ctypes.POINTER(WScene)]),
        "rtw_scene_icosphere": (i32, [u32, u64, ctypes.c_float, ctypes.c_float, u32, ctypes.POINTER(WScene)]),
        "rtw_scene_clusters": (i32, [u32, u64, u32, ctypes.POINTER(WScene)]),
        "rtw_scene_doge": (i32, [ctypes.c_char_p, u32, ctypes.POINTER(WScene)]),
        "rt_obj_load": (i32, [ctypes.c_char_p, u32, ctypes.POINTER(P), ctypes.POINTER(u32)]),
        "rt_obj_last_error": (ctypes.c_char_p, []),
        "rt_free": (None, [P]),
        "rtw_scene_free": (None, [ctypes.POINTER(WScene)]),
        "rtw_rays_camera_tiles": (i32, [P, ctypes.c_float, u32, u32, u32, u32, u32, P]),
        "rtw_rays_shell": (i32, [u32, u64, ctypes.c_float, ctypes.c_float, P]),
        "rtw_rays_box": (i32, [u32, u64, ctypes.c_float, P]),
        "rtw_log_uniform": (u32, [u64, u32, u32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _np_ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


def _check(rc: int, ctx=None, what: str = ""):
    if rc != RT_OK:
        msg = lib().rt_last_error(ctx).decode() if ctx else lib().rt_status_string(rc).decode()
        raise RtError(rc, f"{what}: {msg}" if what else msg)


# ------------------------------------------------------------ host side
class HostScene:
    """Scene arrays in the reference's layouts (numpy)."""

    def __init__(self, triangles, materials, nodes=None, lights=None, spheres=None):
        self.triangles = np.ascontiguousarray(triangles, dtype=TRI_DT)
        self.materials = np.ascontiguousarray(materials, dtype=MAT_DT)
        self.nodes = np.ascontiguousarray(nodes if nodes is not None else np.zeros(0, NODE_DT), dtype=NODE_DT)
        self.lights = np.ascontiguousarray(lights if lights is not None else np.zeros(0, LIGHT_DT), dtype=LIGHT_DT)
        self.spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE_DT), dtype=SPHERE_DT)

    def desc(self) -> SceneDesc:
        return SceneDesc(_np_ptr(self.triangles), len(self.triangles), _np_ptr(self.materials), len(self.materials),
                         _np_ptr(self.nodes), len(self.nodes), _np_ptr(self.lights), len(self.lights),
                         _np_ptr(self.spheres), len(self.spheres))


def _take(ws: WScene) -> HostScene:
    def arr(ptr, n, dt):
        if n == 0:
            return np.zeros(0, dt)
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(ptr)
        # a bytearray keeps every byte (padding included) and stays writable
        return np.frombuffer(bytearray(buf), dtype=dt)
    try:
        return HostScene(arr(ws.triangles, ws.num_triangles, TRI_DT), arr(ws.materials, ws.num_materials, MAT_DT),
                         arr(ws.nodes, ws.num_nodes, NODE_DT), arr(ws.lights, ws.num_lights, LIGHT_DT),
                         arr(ws.spheres, ws.num_spheres, SPHERE_DT))
    finally:
        lib().rtw_scene_free(ctypes.byref(ws))


def pinned_array(shape, dtype) -> np.ndarray:
    """A numpy array in page-locked host memory (rt_host_alloc), freed with the array."""
    import weakref
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(1, count * dt.itemsize)
    ptr = ctypes.c_void_p()
    _check(lib().rt_host_alloc(nbytes, ctypes.byref(ptr)), what="rt_host_alloc")
    buf = (ctypes.c_char * nbytes).from_address(ptr.value)
    arr = np.frombuffer(buf, dtype=np.uint8, count=count * dt.itemsize).view(dt).reshape(shape)
    weakref.finalize(buf, lib().rt_host_free, ptr.value)
    return arr


def scene_soup(n: int, seed: int, threads: int = 8) -> HostScene:
    ws = WScene()
    _check(lib().rtw_scene_soup(n, seed, threads, ctypes.byref(ws)), what="rtw_scene_soup")
    return _take(ws)


def scene_icosphere(level: int, seed: int, amp: float = 0.1, freq: float = 4.0, threads: int = 8) -> HostScene:
    ws = WScene()
    _check(lib().rtw_scene_icosphere(level, seed, amp, freq, threads, ctypes.byref(ws)), what="rtw_scene_icosphere")
    return _take(ws)


def scene_clusters(n: int, seed: int, threads: int = 8) -> HostScene:
    ws = WScene()
    _check(lib().rtw_scene_clusters(n, seed, threads, ctypes.byref(ws)), what="rtw_scene_clusters")
    return _take(ws)


def scene_doge(models_dir: str, threads: int = 8) -> HostScene:
    """The reference's doge scene (RtScene.h:44-101) from its OBJ directory."""
    ws = WScene()
    rc = lib().rtw_scene_doge(models_dir.encode(), threads, ctypes.byref(ws))
    if rc != RT_OK:
        raise RtError(rc, f"rtw_scene_doge({models_dir}): {lib().rt_obj_last_error().decode() or lib().rt_status_string(rc).decode()}")
    return _take(ws)


def obj_load(path: str, material_index: int = 0) -> np.ndarray:
    """Triangle records of one OBJ model (rt_obj_load)."""
    p, n = ctypes.c_void_p(), ctypes.c_uint32()
    rc = lib().rt_obj_load(path.encode(), material_index, ctypes.byref(p), ctypes.byref(n))
    if rc != RT_OK:
        raise RtError(rc, f"rt_obj_load({path}): {lib().rt_obj_last_error().decode() or lib().rt_status_string(rc).decode()}")
    try:
        raw = ctypes.string_at(p.value, n.value * TRI_DT.itemsize) if n.value else b""
        return np.frombuffer(bytearray(raw), TRI_DT).copy()
    finally:
        lib().rt_free(p.value)


def rays_camera_tiles(cam_pos, vfov, w, h, tile, first, count) -> np.ndarray:
    out = np.zeros(count, RAY_DT)
    cam = (ctypes.c_float * 3)(*[float(v) for v in cam_pos])
    _check(lib().rtw_rays_camera_tiles(cam, vfov, w, h, tile, first, count, _np_ptr(out)), what="rtw_rays_camera_tiles")
    return out


def rays_shell(n, seed, radius=3.0, half=0.5) -> np.ndarray:
    out = np.zeros(n, RAY_DT)
    _check(lib().rtw_rays_shell(n, seed, radius, half, _np_ptr(out)), what="rtw_rays_shell")
    return out


def rays_box(n, seed, extent=4.0) -> np.ndarray:
    out = np.zeros(n, RAY_DT)
    _check(lib().rtw_rays_box(n, seed, extent, _np_ptr(out)), what="rtw_rays_box")
    return out


def log_uniform(seed: int, lo: int, hi: int) -> int:
    return int(lib().rtw_log_uniform(seed, lo, hi))


def bvh_build(triangles: np.ndarray, seed: int | None = None, threads: int = 1) -> np.ndarray:
    """Bvh::createBvh.  seed=None consumes this process's rand()."""
    tris = np.ascontiguousarray(triangles, TRI_DT)
    out = np.zeros(max(0, 2 * len(tris) - 1), NODE_DT)
    if seed is None:
        rc = lib().rt_bvh_build(_np_ptr(tris), len(tris), _np_ptr(out))
    else:
        rc = lib().rt_bvh_build_seeded(_np_ptr(tris), len(tris), seed, threads, _np_ptr(out))
    _check(rc, what="rt_bvh_build")
    return out


def scene_check(scene: HostScene):
    """Validates a scene without a device; returns (status, SceneInfo, reason)."""
    info = SceneInfo()
    buf = ctypes.create_string_buffer(512)
    d = scene.desc()
    rc = lib().rt_scene_check(ctypes.byref(d), ctypes.byref(info), buf, 512)
    return rc, info, buf.value.decode()


def scene_lights(triangles: np.ndarray, materials: np.ndarray) -> np.ndarray:
    tris = np.ascontiguousarray(triangles, TRI_DT)
    mats = np.ascontiguousarray(materials, MAT_DT)
    cnt = ctypes.c_uint32(0)
    _check(lib().rt_scene_lights(_np_ptr(tris), len(tris), _np_ptr(mats), len(mats), None, 0, ctypes.byref(cnt)),
           what="rt_scene_lights")
    out = np.zeros(cnt.value, LIGHT_DT)
    _check(lib().rt_scene_lights(_np_ptr(tris), len(tris), _np_ptr(mats), len(mats), _np_ptr(out), cnt.value,
                                 ctypes.byref(cnt)), what="rt_scene_lights")
    return out


# ------------------------------------------------------------ device side
def _dev_ptr(t) -> int:
    return int(t.data_ptr())


def _stream_ptr(stream) -> int | None:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream) or None


def rays_to_device(rays: np.ndarray, device="cuda"):
    import torch
    r = np.ascontiguousarray(rays, RAY_DT)
    return torch.from_numpy(r.view(np.float32).reshape(len(r), 8).copy()).to(device)


def hits_from_device(t) -> np.ndarray:
    a = t.detach().cpu().contiguous().numpy()
    return a.view(np.uint8).reshape(-1).view(HIT_DT).copy()


class Context:
    """An rt_ctx on one device."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        rc = lib().rt_create(ctypes.byref(self._h), device)
        if rc != RT_OK:
            raise RtError(rc, f"rt_create(device={device}) failed: {lib().rt_status_string(rc).decode()}")
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().rt_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what=""):
        _check(rc, self._h, what)

    def set_tuning(self, node_layout=0, refill=0, blocks_per_cu=0, stack_lds=0, tree=0, ray_order=0, steal=0):
        t = Tuning(node_layout, refill, blocks_per_cu, stack_lds, tree, ray_order, steal)
        self.check(lib().rt_tuning_set(self._h, ctypes.byref(t)), "rt_tuning_set")

    def get_tuning(self) -> Tuning:
        t = Tuning()
        self.check(lib().rt_tuning_get(self._h, ctypes.byref(t)), "rt_tuning_get")
        return t

    def upload(self, scene: HostScene) -> "Scene":
        h = ctypes.c_void_p()
        d = scene.desc()
        self.check(lib().rt_scene_upload(self._h, ctypes.byref(d), ctypes.byref(h)), "rt_scene_upload")
        return Scene(self, h)

    def closest_hit(self, scene: "Scene", rays_dev, hits_dev=None, traversal=TRAVERSE_AUTO, max_stack=0, spheres=0,
                    stream=None, n=None):
        import torch
        n = rays_dev.shape[0] if n is None else n
        if hits_dev is None:
            hits_dev = torch.empty((n, 4), dtype=torch.float32, device=rays_dev.device)
        p = QueryParams(traversal, max_stack, spheres, 0)
        self.check(lib().rt_closest_hit(self._h, scene.handle, _dev_ptr(rays_dev), n, _dev_ptr(hits_dev),
                                        ctypes.byref(p), _stream_ptr(stream)), "rt_closest_hit")
        return hits_dev

    def closest_hit_stats(self, scene: "Scene", rays_dev, stream=None):
        import torch
        n = rays_dev.shape[0]
        hits = torch.empty((n, 4), dtype=torch.float32, device=rays_dev.device)
        nv = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        tt = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        self.check(lib().rt_closest_hit_stats(self._h, scene.handle, _dev_ptr(rays_dev), n, _dev_ptr(hits),
                                              _dev_ptr(nv), _dev_ptr(tt), _stream_ptr(stream)), "rt_closest_hit_stats")
        return hits, nv, tt

    def render(self, scene: "Scene", cam_pos, w, h, current_sample=0, vfov=30.0, bounces=2, max_stack=0, spheres=0,
               traversal=TRAVERSE_AUTO, target=None, accum=None, stream=None):
        """One frame (rt_render); returns the RGBA8 target as a [h, w, 4] uint8 tensor."""
        import torch
        dev = f"cuda:{self.device}"
        if target is None:
            target = torch.zeros((h, w, 4), dtype=torch.uint8, device=dev)
        if accum is None:
            accum = torch.zeros((h, w, 4), dtype=torch.uint8, device=dev)
        ubo = make_ubo(cam_pos, current_sample, scene.info)
        prm = RenderParams(vfov, bounces, max_stack, spheres, traversal)
        self.check(lib().rt_render(self._h, scene.handle, ctypes.byref(ubo), ctypes.byref(prm), _dev_ptr(target),
                                   _dev_ptr(accum), w, h, _stream_ptr(stream)), "rt_render")
        return target

    def render_frames(self, scene: "Scene", cam_pos, w, h, frames, first_sample=0, vfov=30.0, bounces=2, max_stack=0,
                      spheres=0, traversal=TRAVERSE_AUTO, target=None, accum=None, stream=None):
        import torch
        dev = f"cuda:{self.device}"
        if target is None:
            target = torch.zeros((h, w, 4), dtype=torch.uint8, device=dev)
        if accum is None:
            accum = torch.zeros((h, w, 4), dtype=torch.uint8, device=dev)
        ubo = make_ubo(cam_pos, first_sample, scene.info)
        prm = RenderParams(vfov, bounces, max_stack, spheres, traversal)
        self.check(lib().rt_render_frames(self._h, scene.handle, ctypes.byref(ubo), ctypes.byref(prm), frames,
                                          _dev_ptr(target), _dev_ptr(accum), w, h, _stream_ptr(stream)),
                   "rt_render_frames")
        return target, accum

    def last_rays(self):
        p, s = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(lib().rt_render_last_rays(self._h, ctypes.byref(p), ctypes.byref(s)), "rt_render_last_rays")
        return p.value, s.value

    def build_bvh_device(self, tris_dev, n: int, nodes_dev, stream=None):
        """rt_bvh_build_device on device buffers (n 48-byte triangles in, 2n-1 48-byte nodes out)."""
        self.check(lib().rt_bvh_build_device(self._h, _dev_ptr(tris_dev), n, _dev_ptr(nodes_dev), _stream_ptr(stream)),
                   "rt_bvh_build_device")

    def build_bvh(self, triangles) -> np.ndarray:
        """The GPU linear BVH of host triangle records, returned as host node records."""
        import torch
        tris = np.ascontiguousarray(triangles, TRI_DT)
        n = len(tris)
        dev = f"cuda:{self.device}"
        td = torch.from_numpy(tris.view(np.uint8).reshape(-1).copy()).to(dev)
        nd = torch.empty(max(1, 2 * n - 1) * NODE_DT.itemsize, dtype=torch.uint8, device=dev)
        self.build_bvh_device(td, n, nd)
        torch.cuda.synchronize(dev)
        return np.frombuffer(bytearray(nd.cpu().numpy().tobytes()), NODE_DT)

    def postprocess_gamma(self, src, dst=None, stream=None):
        """The display pass (rt_postprocess_gamma) of an [h, w, 4] uint8 device image."""
        import torch
        h, w = src.shape[0], src.shape[1]
        if dst is None:
            dst = torch.empty_like(src)
        self.check(lib().rt_postprocess_gamma(self._h, _dev_ptr(src), _dev_ptr(dst), w, h, _stream_ptr(stream)),
                   "rt_postprocess_gamma")
        return dst

    def set_render_timing(self, enable: bool):
        self.check(lib().rt_render_set_timing(self._h, 1 if enable else 0), "rt_render_set_timing")

    def last_query_stats(self) -> QueryStats:
        st = QueryStats()
        self.check(lib().rt_query_last_stats(self._h, ctypes.byref(st)), "rt_query_last_stats")
        return st

    def last_render_stats(self) -> RenderStats:
        st = RenderStats()
        self.check(lib().rt_render_last_stats(self._h, ctypes.byref(st)), "rt_render_last_stats")
        return st

    def set_render_counting(self, enable: bool):
        self.check(lib().rt_render_set_counting(self._h, 1 if enable else 0), "rt_render_set_counting")

    def last_render_counts(self) -> tuple[int, int]:
        """(node records, triangle records) the last counted render's rays read."""
        nv, nt = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(lib().rt_render_last_counts(self._h, ctypes.byref(nv), ctypes.byref(nt)), "rt_render_last_counts")
        return int(nv.value), int(nt.value)


class Batch:
    """An rt_batch: items (scene, rays, hits) traced by one persistent launch."""

    def __init__(self, ctx: "Context", items, config: BatchConfig | None = None):
        self.ctx = ctx
        self._keep = list(items)
        arr = (BatchItem * max(1, len(items)))()
        for k, (scene, rays, hits) in enumerate(items):
            arr[k] = BatchItem(scene.handle.value if hasattr(scene.handle, "value") else scene.handle,
                               _dev_ptr(rays) if rays.numel() else None, _dev_ptr(hits) if hits.numel() else None,
                               rays.shape[0], 0)
        h = ctypes.c_void_p()
        ctx.check(lib().rt_batch_create(ctx.handle, arr, len(items), ctypes.byref(config) if config else None,
                                        ctypes.byref(h)), "rt_batch_create")
        self.handle = h

    def launch(self, stream=None):
        self.ctx.check(lib().rt_batch_launch(self.ctx.handle, self.handle, _stream_ptr(stream)), "rt_batch_launch")

    def set_config(self, config: BatchConfig):
        self.ctx.check(lib().rt_batch_set_config(self.ctx.handle, self.handle, ctypes.byref(config)),
                       "rt_batch_set_config")

    def config(self) -> BatchConfig:
        c = BatchConfig()
        _check(lib().rt_batch_get_config(self.handle, ctypes.byref(c)), what="rt_batch_get_config")
        return c

    def autotune(self, trials=3, cache_path: str | None = None) -> BatchConfig:
        c = BatchConfig()
        self.ctx.check(lib().rt_batch_autotune(self.ctx.handle, self.handle, trials,
                                               cache_path.encode() if cache_path else None, ctypes.byref(c)),
                       "rt_batch_autotune")
        return c

    def destroy(self):
        if self.handle:
            self.ctx.check(lib().rt_batch_destroy(self.ctx.handle, self.handle), "rt_batch_destroy")
            self.handle = None


class FrameLoop:
    """An rt_frame_loop: the reference's per-frame sequence with its own
    target, accumulation and display images on the device."""

    def __init__(self, ctx: "Context", scene: "Scene", w: int, h: int, vfov=30.0, bounces=2, max_stack=0,
                 spheres=0, traversal=TRAVERSE_AUTO):
        self.ctx, self.scene, self.w, self.h = ctx, scene, w, h
        prm = RenderParams(vfov, bounces, max_stack, spheres, traversal)
        hnd = ctypes.c_void_p()
        ctx.check(lib().rt_frame_loop_create(ctx.handle, scene.handle, ctypes.byref(prm), w, h, ctypes.byref(hnd)),
                  "rt_frame_loop_create")
        self.handle = hnd

    def step(self, cam_pos, time=0.0, moved=False, stream=None):
        cam = (ctypes.c_float * 3)(*[float(v) for v in cam_pos])
        self.ctx.check(lib().rt_frame_loop_step(self.handle, cam, float(time), 1 if moved else 0, _stream_ptr(stream)),
                       "rt_frame_loop_step")

    def display(self, stream=None):
        self.ctx.check(lib().rt_frame_loop_display(self.handle, _stream_ptr(stream)), "rt_frame_loop_display")

    def _images(self):
        t, a, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        cs = ctypes.c_uint32()
        _check(lib().rt_frame_loop_images(self.handle, ctypes.byref(t), ctypes.byref(a), ctypes.byref(d),
                                          ctypes.byref(cs)), what="rt_frame_loop_images")
        return t.value, a.value, d.value, cs.value

    @property
    def current_sample(self) -> int:
        return self._images()[3]

    def image(self, which: str = "target", stream=None) -> np.ndarray:
        """Copies one of the loop's images ("target", "accum", "display") to the host."""
        out = np.empty((self.h, self.w, 4), np.uint8)
        k = {"target": 0, "accum": 1, "display": 2}[which]
        self.ctx.check(lib().rt_frame_loop_read(self.handle, k, out.ctypes.data, _stream_ptr(stream)),
                       "rt_frame_loop_read")
        return out

    def close(self):
        if self.handle:
            lib().rt_frame_loop_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gamma_table() -> np.ndarray:
    out = np.zeros(256, np.uint8)
    _check(lib().rt_gamma_table(out.ctypes.data), what="rt_gamma_table")
    return out


def image_write(path: str, rgba: np.ndarray):
    """rt_image_write of an [h, w, 4] uint8 host image (.png or .ppm)."""
    img = np.ascontiguousarray(rgba, np.uint8)
    h, w, c = img.shape
    if c != 4:
        raise ValueError("image_write: expects [h, w, 4] RGBA8")
    _check(lib().rt_image_write(path.encode(), img.ctypes.data, w, h), what=f"rt_image_write({path})")


class Server:
    """An rt_server: resident scenes and a request loop on its own thread."""

    def __init__(self, ctx: "Context", mode=SERVE_BATCHED, slots=0, max_batch_rays=0, max_render_pixels=0,
                 hbm_budget=0, deadline_ms=0):
        self.ctx = ctx
        cfg = ServerConfig(mode, slots, max_batch_rays, max_render_pixels, hbm_budget, deadline_ms)
        h = ctypes.c_void_p()
        ctx.check(lib().rt_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)), "rt_server_create")
        self.handle = h
        self._keep = {}

    def _check(self, rc, what):
        if rc != RT_OK:
            raise RtError(rc, f"{what}: {lib().rt_server_last_error(self.handle).decode()}")

    def add_scene(self, hs: HostScene) -> int:
        d = hs.desc()
        sid = ctypes.c_uint64()
        self._check(lib().rt_server_scene_add(self.handle, ctypes.byref(d), ctypes.byref(sid)), "rt_server_scene_add")
        return sid.value

    def remove_scene(self, sid: int):
        self._check(lib().rt_server_scene_remove(self.handle, sid), "rt_server_scene_remove")

    def scenes(self):
        cnt = ctypes.c_uint32()
        lib().rt_server_scene_list(self.handle, None, 0, ctypes.byref(cnt))
        ids = (ctypes.c_uint64 * max(1, cnt.value))()
        lib().rt_server_scene_list(self.handle, ids, cnt.value, ctypes.byref(cnt))
        return [ids[i] for i in range(cnt.value)]

    def submit_query(self, sid: int, rays: np.ndarray, hits: np.ndarray) -> int:
        rid = ctypes.c_uint64()
        self._check(lib().rt_server_submit_query(self.handle, sid, _np_ptr(rays), len(rays), _np_ptr(hits),
                                                 ctypes.byref(rid)), "rt_server_submit_query")
        self._keep[rid.value] = (rays, hits)
        return rid.value

    def submit_render(self, sid: int, cam_pos, w, h, target: np.ndarray, current_sample=0, vfov=60.0, bounces=4):
        ubo = make_ubo(cam_pos, current_sample)
        prm = RenderParams(vfov, bounces, 0, 0, 0)
        rid = ctypes.c_uint64()
        self._check(lib().rt_server_submit_render(self.handle, sid, ctypes.byref(ubo), ctypes.byref(prm), w, h,
                                                  target.ctypes.data, ctypes.byref(rid)), "rt_server_submit_render")
        self._keep[rid.value] = (target,)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> int:
        st = ctypes.c_int()
        self._check(lib().rt_server_wait(self.handle, rid, timeout_ms, ctypes.byref(st)), "rt_server_wait")
        self._keep.pop(rid, None)
        return st.value

    def cancel(self, rid: int) -> int:
        """Cancels a queued request; returns the status (RT_OK, RT_ERR_BUSY
        when it is already running, RT_ERR_NOT_FOUND when unknown or done)."""
        return lib().rt_server_cancel(self.handle, rid)

    def drain(self):
        self._check(lib().rt_server_drain(self.handle), "rt_server_drain")

    def stats(self) -> ServerStats:
        s = ServerStats()
        self._check(lib().rt_server_stats_get(self.handle, ctypes.byref(s)), "rt_server_stats_get")
        return s

    def reset_stats(self):
        lib().rt_server_stats_reset(self.handle)

    def poisson(self, rate, duration, render_fraction=0.1, swap_interval=10.0, query_min=1000, query_max=262144,
                render_w=512, render_h=512, render_bounces=4, swap_triangles=500000, seed=7, threads=16) -> ServeResult:
        p = ServeParams(rate, duration, render_fraction, swap_interval, query_min, query_max, render_w, render_h,
                        render_bounces, swap_triangles, seed, threads, 0)
        r = ServeResult()
        self._check(lib().rtw_serve_poisson(self.handle, ctypes.byref(p), ctypes.byref(r)), "rtw_serve_poisson")
        return r

    def poisson_sampled(self, rate, duration, every, capacity, swaps_capacity=16, render_fraction=0.1,
                        swap_interval=10.0, query_min=1000, query_max=262144, render_w=512, render_h=512,
                        render_bounces=4, swap_triangles=500000, seed=7, threads=16):
        """A Poisson run that keeps every `every`-th arrival's result (up to
        `capacity`).  Returns (result, samples, swap_ids): each sample is a
        dict with kind ("query"/"render"), scene_id, arrival, count and its
        hits (HIT_DT records) or RGBA8 image; swap_ids[k] is the server id
        of rtw_scene_clusters(swap_triangles, seed + 1000 + k) (0 if that
        swap failed)."""
        p = ServeParams(rate, duration, render_fraction, swap_interval, query_min, query_max, render_w, render_h,
                        render_bounces, swap_triangles, seed, threads, 0)
        hits = [np.zeros(query_max, HIT_DT) for _ in range(capacity)]
        imgs = [np.zeros((render_h, render_w, 4), np.uint8) for _ in range(capacity)]
        arr = (ServeSample * max(1, capacity))()
        for i in range(capacity):
            arr[i].hits = hits[i].ctypes.data
            arr[i].image = imgs[i].ctypes.data
        ids = (ctypes.c_uint64 * max(1, swaps_capacity))()
        smp = ServeSampling(every, capacity, arr, 0, swaps_capacity, ids, 0, 0)
        r = ServeResult()
        self._check(lib().rtw_serve_poisson_sampled(self.handle, ctypes.byref(p), ctypes.byref(smp), ctypes.byref(r)),
                    "rtw_serve_poisson_sampled")
        samples = []
        for i in range(smp.taken):
            e = arr[i]
            d = {"kind": "render" if e.kind else "query", "scene_id": int(e.scene_id), "arrival": int(e.arrival),
                 "count": int(e.count)}
            if e.kind:
                d["image"] = imgs[i]
            else:
                d["hits"] = hits[i][:e.count]
            samples.append(d)
        return r, samples, [int(ids[k]) for k in range(smp.swaps_taken)]

    def close(self):
        if self.handle:
            lib().rt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle
        info = SceneInfo()
        _check(lib().rt_scene_info_get(handle, ctypes.byref(info)), what="rt_scene_info_get")
        self.info = info

    def release(self):
        if self.handle:
            self.ctx.check(lib().rt_scene_release(self.ctx.handle, self.handle), "rt_scene_release")
            self.handle = None


def make_ubo(cam_pos, current_sample=0, info=None, time=0.0) -> Ubo:
    u = Ubo()
    for k in range(3):
        u.camPos[k] = float(cam_pos[k])
    u.time = time
    u.currentSample = current_sample
    if info is not None:
        u.numTriangles, u.numLights, u.numSpheres = info.num_triangles, info.num_lights, info.num_spheres
    return u
