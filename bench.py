#!/usr/bin/env python3
"""Benchmark of the ray-traversal hot path on one MI355X.

Step = one rt_closest_hit over the whole ray batch (the plain kernel) on the
clustered-soup scene: 4,194,304 triangles in 4096 Gaussian clusters (centres
U[-4,4]^3, sigma 0.15, edges U[0.005,0.03]), BVH from the reference builder,
and 4,194,304 incoherent rays (origins U[-4,4]^3, directions uniform on the
sphere), all resident in HBM before the timed region.

Prints one JSON line: throughput in Mrays/s (one ray = one hit_bvh
traversal), the roofline of the traversal kernel (algorithmic bytes = node
records x 64 B (4-wide; 32 B with --layout q16) + triangle records x 48 B + ray 32 B +
hit 16 B, counted per ray by an untimed pass of the counting kernel, over
the traversal kernel's own time from stage events on its stream; `value`
counts the whole call, the ray sort included), the
CPU baseline (the
oracle's restatement of hit_bvh in the shader's visiting order, on a bounded
sample of the same rays), and side measurements:
  render   the 1920x1080 x 4 spp, 8-bounce render of the same scene;
  device_build  the GPU linear-BVH builder on the same triangles, the host
           builder's time beside it, and the query on the device-built tree;
  small_render, query_medium  the 1,024-triangle soup render and the
           81,920-triangle icosphere query;
  doge_frame_loop  the reference application's own frame: its doge scene at
           800x600, two bounces, render + accumulation copy + display pass;
  batched  64 clustered-soup scenes (triangles log-uniform in [1k, 1M], seeds
           100+i) with ray batches log-uniform in [4k, 2M] (seeds 200+i):
           64 plain launches against one autotuned persistent launch
           replayed from a hipGraph;
  served   16 resident scenes (log-uniform [10k, 2M], seeds 300-315) under
           an HBM budget of 1.05x their footprint, Poisson(4000/s) requests
           (10% 512x512 1-spp 4-bounce renders, the rest queries of
           log-uniform [1k, 256k] rays), a new 500k-triangle scene every 10 s:
           p50/p99 latency and Mrays/s of the batching server, against the
           plain kernel serving the first seconds of the same trace FIFO.
Progress lines go to stderr; the JSON line is the only stdout.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import csv
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "vulkan-compute-ray-tracing_amd")]

METRIC = "incoherent Mrays/s + p50/p99 latency; batched/served vs plain kernel, 1 GPU"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_packed_clusters4M_traffic.txt")
CANON_SAMPLE = 65536


# Host threads of the CPU baseline: the cores this process may run on,
# capped at the 16-core host share a GPU has on the MI355X boxes it runs on.
CPU_SHARE = 16
# Shader engines of the MI355X (8 XCDs x 4): SQ_BUSY_CYCLES sums over them.
SHADER_ENGINES = 32


def visible_cores() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def host_threads() -> int:
    return max(1, min(CPU_SHARE, visible_cores()))


def kernel_sources_sha() -> str:
    """Hash of the traversal kernel's sources, to tell whether a committed
    counter figure was taken on the kernel this tree builds."""
    import hashlib
    h = hashlib.sha256()
    for f in ("rt_trace_dev.hpp", "rt_trace.hip"):
        with open(os.path.join(ROOT, "vulkan-compute-ray-tracing_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def read_traffic():
    """L2-to-memory bytes per launch from the committed counter profile, and
    whether that profile was taken on the kernel sources of this tree."""
    vals = {}
    try:
        for line in open(TRAFFIC_FILE):
            if "=" in line and not line.startswith("#"):
                k, v = line.strip().split("=", 1)
                vals[k] = v
    except OSError:
        return None, None
    if "bytes_per_launch" not in vals:
        return None, None
    return float(vals["bytes_per_launch"]), vals.get("kernel_sources_sha") == kernel_sources_sha()


# Counter passes (rocprofv3 takes at most 4 TCC and 8 SQ counters per run).
# L2-to-memory bytes: every 32-byte unit a TCC read or write request sends
# to DRAM (a 64-byte request counts 2, a 128-byte one 4) -- what leaves the
# L2 for HBM, which the Infinity Cache may still serve.
PMC_PASSES = (
    ("TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum",
     "TCC_HIT_sum"),
    ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_BUSY_CYCLES", "SQ_ACTIVE_INST_VALU",
     "SQ_THREAD_CYCLES_VALU", "TCC_MISS_sum"),
)
# kernel classes counted: the timed query kernel (not the counting variant),
# the render's per-bounce traversals and its shading kernel
KERNEL_CLASSES = (("query", re.compile(r"k_trace_packed<\d+, \d+, \d+, false|k_trace_packedILi\d+ELi\d+ELi\d+ELb0")),
                  ("render_trace", re.compile(r"k_trace_render")), ("shade", re.compile(r"k_shade")),
                  ("copy", re.compile(r"copyBuffer")))
# the known-size copy the counter workload ends with: it calibrates the
# byte counts (bytes read and written by a plain device-to-device copy)
CALIBRATION_MIB = 1024


def add_counter_rows(agg: dict, rows, k: int):
    """Adds one counter pass (rows of rocprofv3's counter_collection.csv) to
    agg[class]: counters summed over the class's dispatches, the dispatch
    count of pass k, and (pass 0) each dispatch's duration."""
    # per dispatch: its kernel class, grid size and counter values
    per = {}
    for row in rows:
        name = row.get("Kernel_Name", "")
        cls = next((c for c, rx in KERNEL_CLASSES if rx.search(name)), None)
        if cls is None:
            continue
        key = row.get("Dispatch_Id") or row.get("Correlation_Id")
        dd = per.setdefault((cls, key), {"grid": float(row.get("Grid_Size", 0) or 0), "c": {}, "ns": None})
        dd["c"][row.get("Counter_Name", "")] = float(row.get("Counter_Value", 0) or 0)
        if row.get("Start_Timestamp") and row.get("End_Timestamp"):
            dd["ns"] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
    # the calibration copy is the class's largest dispatch (other copies are small transfers)
    big_copy = max((key for (cls, key) in per if cls == "copy"), key=lambda kk: per[("copy", kk)]["grid"],
                   default=None)
    for (cls, key), dd in per.items():
        if cls == "copy" and key != big_copy:
            continue
        a = agg.setdefault(cls, {})
        for ctr, v in dd["c"].items():
            a[ctr] = a.get(ctr, 0.0) + v
        a[f"_dispatches{k}"] = a.get(f"_dispatches{k}", 0) + 1
        if k == 0 and dd["ns"] is not None:
            a.setdefault("_ns", {})[key] = dd["ns"]


def run_counter_passes(args) -> dict | None:
    """Runs this script's counter workload (--pmc-child: the timed query,
    then the render unless --no-render) under rocprofv3 once per counter
    pass, and sums each counter over each kernel class's dispatches."""
    if not shutil.which("rocprofv3"):
        log("counters: rocprofv3 not found")
        return None
    agg = {}
    for k, counters in enumerate(PMC_PASSES):
        d = tempfile.mkdtemp(prefix="rtamd_pmc_")
        # a pass takes a few seconds; a run that exceeds the limit is killed and its pass reads as absent
        cmd = ["timeout", "-s", "KILL", "120", "rocprofv3", "--pmc", *counters, "-f", "csv", "-d", d, "-o", "pmc",
               "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--tris", str(args.tris),
               "--rays", str(args.rays), "--layout", args.layout, "--refill", str(args.refill)]
        if args.no_render:
            cmd.append("--no-render")
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        rows = []
        for f in files:
            with open(f, newline="") as fh:
                rows.extend(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        # rocprofv3 may end with a signal after writing whole results: rows decide
        if not rows:
            log(f"counters: pass {k} wrote no rows (exit {r.returncode}): {r.stderr.strip()[-300:]}")
            return None
        add_counter_rows(agg, rows, k)
        log(f"counters: pass {k} ({len(counters)} counters) in {time.perf_counter() - t0:.1f} s")
    return agg


def counter_summary(c: dict, num_cus: int) -> dict:
    """Per-dispatch means of one kernel class and the figures they give."""
    d0, d1 = c.get("_dispatches0", 0), c.get("_dispatches1", 0)
    if not d0 or not d1:
        return {}
    g = lambda k, d: c.get(k, 0.0) / d
    rd = 32.0 * g("TCC_EA0_RDREQ_DRAM_32B_sum", d0)
    wr = 32.0 * (g("TCC_EA0_WRREQ_WRITE_DRAM_32B_sum", d0) + g("TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum", d0))
    hit, miss = g("TCC_HIT_sum", d0), g("TCC_MISS_sum", d1)
    act, busy = g("SQ_ACTIVE_INST_VALU", d1), g("SQ_BUSY_CYCLES", d1)
    simds = 4 * num_cus
    out = {"dispatches": d0, "dram_read_bytes": rd, "dram_write_bytes": wr, "dram_bytes": rd + wr,
           "l2_hit": hit / (hit + miss) if hit + miss else None,
           "valu_insts": g("SQ_INSTS_VALU", d1), "salu_insts": g("SQ_INSTS_SALU", d1), "waves": g("SQ_WAVES", d1),
           # a wave64 VALU instruction holds a 16-lane SIMD for 4 cycles; SQ_BUSY_CYCLES sums the shader engines
           "valu_busy": act * 4.0 / (simds * busy / SHADER_ENGINES) if busy else None,
           "valu_lanes_active": g("SQ_THREAD_CYCLES_VALU", d1) / (64.0 * act) if act else None}
    ns = c.get("_ns", {})
    if ns:
        out["profiled_kernel_ms"] = sum(ns.values()) / len(ns) / 1e6
    return out


def pmc_child(args, threads: int):
    """The workload the counter passes profile: the timed query (one warm-up
    and two launches) and, unless --no-render, one 1080p 4-frame render."""
    import torch
    import rtamd
    ctx = rtamd.Context(0)
    layout = {"w4": rtamd.LAYOUT_W4, "q16": rtamd.LAYOUT_Q16, "f32": rtamd.LAYOUT_F32}[args.layout]
    ctx.set_tuning(node_layout=layout, refill=args.refill)
    hs = rtamd.scene_clusters(args.tris, 3, threads=threads)
    scene = ctx.upload(hs)
    rays = rtamd.rays_to_device(rtamd.rays_box(args.rays, 4, 4.0))
    hits = torch.empty((rays.shape[0], 4), dtype=torch.float32, device="cuda")
    for _ in range(3):
        ctx.closest_hit(scene, rays, hits)
    torch.cuda.synchronize()
    if not args.no_render:
        tgt = torch.zeros((1080, 1920, 4), dtype=torch.uint8, device="cuda")
        acc = torch.zeros_like(tgt)
        ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
        torch.cuda.synchronize()
    scene.release()
    ctx.close()
    src = torch.ones(CALIBRATION_MIB << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def median_ms(fn, reps: int) -> float:
    import torch
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps):
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    return sorted(ts)[len(ts) // 2]


def bench_batched(ctx, threads: int, reps: int):
    """64 scenes x uneven ray batches: plain launches vs one batched launch."""
    import torch
    import rtamd
    t0 = time.perf_counter()
    scenes, plain_items, batch_items = [], [], []
    for i in range(64):
        hs = rtamd.scene_clusters(rtamd.log_uniform(100 + i, 1000, 1000000), 100 + i, threads=threads)
        sc = ctx.upload(hs)
        c = rtamd.log_uniform(200 + i, 4096, 2097152)
        rays = rtamd.rays_to_device(rtamd.rays_box(c, 200 + i, 4.0))
        scenes.append(sc)
        plain_items.append((sc, rays, torch.empty((c, 4), dtype=torch.float32, device="cuda")))
        batch_items.append((sc, rays, torch.empty((c, 4), dtype=torch.float32, device="cuda")))
    tris = sum(s.info.num_triangles for s in scenes)
    nrays = sum(r.shape[0] for _, r, _ in plain_items)
    log(f"batched: 64 scenes, {tris} triangles, {nrays} rays built in {time.perf_counter() - t0:.1f} s")

    def plain():
        for sc, r, h in plain_items:
            ctx.closest_hit(sc, r, h)

    t_plain = median_ms(plain, reps)
    b = rtamd.Batch(ctx, batch_items)
    t0 = time.perf_counter()
    best = b.autotune(trials=3)
    t_tune = time.perf_counter() - t0
    t_batch = median_ms(b.launch, reps)
    torch.cuda.synchronize()
    same = all(torch.equal(hp.view(torch.int32), hb.view(torch.int32))
               for (_, _, hp), (_, _, hb) in zip(plain_items, batch_items))
    log(f"batched: plain {t_plain:.2f} ms, batched {t_batch:.2f} ms, ratio {t_plain / t_batch:.2f}")
    out = {"workload": "64 clustered-soup scenes, triangles log-uniform [1k,1M] (seeds 100+i), ray batches "
                       "log-uniform [4k,2M] (seeds 200+i, origins U[-4,4]^3, uniform directions)",
           "scenes": 64, "triangles": tris, "rays": nrays,
           "plain_ms": round(t_plain, 3), "plain_mrays_per_s": round(nrays / t_plain / 1e3, 1),
           "batched_ms": round(t_batch, 3), "mrays_per_s": round(nrays / t_batch / 1e3, 1),
           "ratio": round(t_plain / t_batch, 3), "hits_identical": bool(same),
           "autotune_s": round(t_tune, 1), "config": best.as_dict()}
    b.destroy()
    del b, plain_items, batch_items, scenes
    torch.cuda.synchronize()
    return out


def bench_served(ctx, threads: int, seconds: float, fifo_seconds: float, rates=()):
    """Poisson request stream against the batching server and the FIFO plain server."""
    import rtamd
    t0 = time.perf_counter()
    hss = [rtamd.scene_clusters(rtamd.log_uniform(300 + i, 10000, 2000000), 300 + i, threads=threads)
           for i in range(16)]
    tris = sum(len(h.triangles) for h in hss)
    log(f"served: 16 scenes, {tris} triangles built in {time.perf_counter() - t0:.1f} s")
    probe = rtamd.Server(ctx)
    for h in hss:
        probe.add_scene(h)
    budget = int(probe.stats().resident_bytes * 1.05)
    probe.close()

    def run(mode, slots, duration, warm):
        srv = rtamd.Server(ctx, mode=mode, slots=slots, hbm_budget=budget)
        for h in hss:
            srv.add_scene(h)
        if warm:
            srv.poisson(rate=4000.0, duration=1.0, swap_interval=0.0, seed=11, threads=threads)
        r = srv.poisson(rate=4000.0, duration=duration, render_fraction=0.1, swap_interval=10.0,
                        swap_triangles=500000, seed=7, threads=threads)
        srv.close()
        st = r.stats
        return {"requests": int(r.requests), "queries": int(r.queries), "renders": int(r.renders),
                "completed": int(st.completed), "failed": int(st.failed), "launches": int(st.launches),
                "scenes_added": int(st.scenes_added), "scenes_evicted": int(st.scenes_evicted),
                "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "mean_ms": round(st.mean_ms, 3),
                "max_ms": round(st.max_ms, 3), "rays": int(st.rays), "wall_s": round(r.wall_s, 2),
                "mrays_per_s": round(st.rays / r.wall_s / 1e6, 1)}

    served = run(rtamd.SERVE_BATCHED, 0, seconds, True)
    log(f"served: p50 {served['p50_ms']} ms p99 {served['p99_ms']} ms, {served['mrays_per_s']} Mrays/s")
    fifo = run(rtamd.SERVE_FIFO, 1, fifo_seconds, False)
    log(f"served: fifo p50 {fifo['p50_ms']} ms, {fifo['mrays_per_s']} Mrays/s")
    capacity = []
    for rate in rates:
        srv = rtamd.Server(ctx, hbm_budget=budget)
        for h in hss:
            srv.add_scene(h)
        r = srv.poisson(rate=float(rate), duration=6.0, render_fraction=0.1, swap_interval=3.0,
                        swap_triangles=500000, seed=7, threads=threads)
        srv.close()
        st = r.stats
        capacity.append({"rate_per_s": rate, "requests": int(r.requests), "p50_ms": round(st.p50_ms, 3),
                         "p99_ms": round(st.p99_ms, 3), "wall_s": round(r.wall_s, 2),
                         "mrays_per_s": round(st.rays / r.wall_s / 1e6, 1)})
        log(f"served: {rate}/s p99 {capacity[-1]['p99_ms']} ms, {capacity[-1]['mrays_per_s']} Mrays/s")
    out = {"workload": "16 resident clustered-soup scenes, triangles log-uniform [10k,2M] (seeds 300-315), "
                       "HBM budget 1.05x their footprint, Poisson(4000/s) arrivals for "
                       f"{seconds:g} s: 10% 512x512 1-spp 4-bounce renders, 90% closest-hit queries of "
                       "log-uniform [1k,256k] rays, a new 500k-triangle scene every 10 s (LRU eviction)",
           "triangles": tris, "hbm_budget_bytes": budget, **served,
           "fifo": {"sample": f"first {fifo_seconds:g} s of the same arrival sequence, plain kernel one "
                              f"request at a time", **fifo},
           "ratio": round(served["mrays_per_s"] / max(fifo["mrays_per_s"], 1e-9), 3)}
    if capacity:
        out["capacity"] = capacity
    return out


def lavapipe_probe() -> str:
    """Whether the reference shader could run on Mesa's CPU Vulkan driver here."""
    import ctypes.util
    import glob
    import shutil
    missing = []
    if not ctypes.util.find_library("vulkan"):
        missing.append("libvulkan.so.1")
    icds = [f for d in ("/usr/share/vulkan/icd.d", "/etc/vulkan/icd.d", "/usr/local/share/vulkan/icd.d")
            for f in glob.glob(f"{d}/lvp_icd*.json")]
    if not icds:
        missing.append("lvp_icd*.json")
    if not (shutil.which("glslangValidator") or shutil.which("glslc")):
        missing.append("glslangValidator/glslc")
    if missing:
        return "absent (missing: " + ", ".join(missing) + ")"
    return "present (not used: no Vulkan host in this build)"


def cpu_baseline(hs, rays_np, seconds: float, threads: int):
    """The oracle (CPU restatement of the shader) on the host cores: the
    4M-triangle query on a calibrated sample (one warm-up, median of 5), and
    beside it the small render and the icosphere query."""
    import numpy as np
    import oracle as O
    import rtamd
    osc = O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)
    per_run = seconds / 5.0
    m = 64 * threads
    for _ in range(8):                          # calibration; the last run is the warm-up
        t0 = time.perf_counter()
        O.closest_hit(osc, rays_np[:m], O.TRAV_REFERENCE, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.6 * per_run or m == len(rays_np):
            break
        m = int(min(len(rays_np), max(2 * m, m * per_run / max(dt, 1e-3))))
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.closest_hit(osc, rays_np[:m], O.TRAV_REFERENCE, threads=threads)
        runs.append(time.perf_counter() - t0)
    dt = sorted(runs)[2]
    k = min(CANON_SAMPLE, len(rays_np))
    _, nv, tt, _ = O.closest_hit(osc, rays_np[:k], O.TRAV_ORDERED, stats=True, threads=threads)

    # the reference's small render and the icosphere query, same inputs as the GPU side
    soup = rtamd.scene_soup(1024, 1, threads=threads)
    t0 = time.perf_counter()
    _, small_rays = O.render(O.Scene(soup.triangles, soup.materials, soup.nodes, soup.lights, soup.spheres),
                             [0.0, 3.0, 0.0], 256, 256, vfov=60.0, bounces=4, threads=threads)
    small_s = time.perf_counter() - t0
    ico = rtamd.scene_icosphere(6, 2, 0.1, 4.0, threads=threads)
    isc = O.Scene(ico.triangles, ico.materials, ico.nodes, ico.lights, ico.spheres)
    coh = rtamd.rays_camera_tiles([0, 3, 0], 60.0, 1024, 1024, 8, 0, 524288)[::8]   # spread over the half image
    inc = rtamd.rays_shell(65536, 2024, 3.0, 0.5)
    medium = {}
    for name, rr in (("coherent", coh), ("incoherent", inc)):
        t0 = time.perf_counter()
        O.closest_hit(isc, rr, O.TRAV_REFERENCE, threads=threads)
        medium[f"{name}_mrays_per_s"] = round(len(rr) / (time.perf_counter() - t0) / 1e6, 4)
    return {"value": round(m / dt / 1e6, 6), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "host_cores_visible": visible_cores(),
            "cores_note": f"threads = the cores this process may use, capped at the {CPU_SHARE}-core host share "
                          "one GPU has on this machine",
            "sample": f"first {m} of the {len(rays_np)} incoherent rays, oracle hit_bvh in the shader's "
                      f"visiting order (line test, both children pushed), {threads} threads, median of 5 runs "
                      f"of {dt:.1f} s after a warm-up; CPU restatement, not lavapipe",
            "cpu_model": cpu_model(), "lavapipe": lavapipe_probe(),
            "small_render": {"rays": int(small_rays), "s": round(small_s, 3),
                             "mrays_per_s": round(small_rays / small_s / 1e6, 4),
                             "sample": "the whole 256x256, 4-bounce render of the 1024-triangle soup"},
            "query_medium": dict(medium, sample="every 8th of the 524288 coherent camera rays and the first "
                                                "65536 incoherent shell rays on the level-6 icosphere"),
            "canonical": {"rays": k, "traversal": "oracle ordered, t-culled (nearer child first)",
                          "node_visits_per_ray": round(float(nv.mean()), 3),
                          "tri_tests_per_ray": round(float(tt.mean()), 3)}}


def bench_device_build(ctx, hs, rays, threads: int, layout: int, node_bytes: int):
    """GPU linear-BVH build of the query scene, the host builder beside it,
    and the query on the device-built tree."""
    import numpy as np
    import torch
    import rtamd
    n = len(hs.triangles)
    td = torch.from_numpy(np.ascontiguousarray(hs.triangles).view(np.uint8).reshape(-1).copy()).cuda()
    nd = torch.empty((2 * n - 1) * 48, dtype=torch.uint8, device="cuda")
    ms = median_ms(lambda: ctx.build_bvh_device(td, n, nd), 5)
    t0 = time.perf_counter()
    rtamd.bvh_build(hs.triangles, seed=1, threads=threads)
    host_s = time.perf_counter() - t0
    nodes = np.frombuffer(bytearray(nd.cpu().numpy().tobytes()), rtamd.NODE_DT)
    dev = ctx.upload(rtamd.HostScene(hs.triangles, hs.materials, nodes, hs.lights, hs.spheres))
    _, nv, tt = ctx.closest_hit_stats(dev, rays)
    v_node = float(nv.to(torch.float64).mean().item())
    v_tri = float(tt.to(torch.float64).mean().item())
    hits = torch.empty((rays.shape[0], 4), dtype=torch.float32, device="cuda")
    tq = median_ms(lambda: ctx.closest_hit(dev, rays, hits), 5)
    bpr = v_node * node_bytes + v_tri * 48 + 48
    out = {"workload": f"linear BVH (63-bit Morton codes, radix tree) of the {n}-triangle query scene, then the "
                       "same incoherent query on that tree",
           "build_ms": round(ms, 3), "host_builder_s": round(host_s, 3), "host_threads": threads,
           "fast_path": int(dev.info.fast_path), "tree_height": int(dev.info.tree_height),
           "query_ms": round(tq, 4), "mrays_per_s": round(rays.shape[0] / tq / 1e3, 1),
           "node_records_per_ray": round(v_node, 2), "tri_records_per_ray": round(v_tri, 3),
           "achieved_gbs": round(bpr * rays.shape[0] / (tq / 1e3) / 1e9, 1)}
    log(f"device build: {ms:.2f} ms (host {host_s:.2f} s), query on it {out['mrays_per_s']} Mrays/s")
    del td, nd, dev, hits
    return out


def cpu_model() -> str:
    """The first processor's identity as /proc/cpuinfo gives it: model name,
    family / model / stepping numbers and clock (a virtual machine may report
    a shortened model name, so the numbers are kept beside it)."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                break
            if ":" in line:
                k, v = line.split(":", 1)
                fields.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    parts = [fields.get("model name", "unknown")]
    ids = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if ids:
        parts.append(", ".join(ids))
    if "cpu MHz" in fields:
        parts.append(f"{float(fields['cpu MHz']):.0f} MHz")
    return "; ".join(parts)


def copy_bandwidth(mib: int = 1024) -> float:
    """Device-to-device copy of `mib` MiB: (read + written bytes) / time, GB/s."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms = median_ms(lambda: dst.copy_(src), 5)
    del src, dst
    return 2.0 * (mib << 20) / (ms / 1e3) / 1e9


def bench_small_and_medium(ctx, threads: int):
    """The reference's own small render (1,024-triangle soup) and the
    81,920-triangle icosphere query (coherent and incoherent halves)."""
    import torch
    import rtamd
    out = {}
    hs = rtamd.scene_soup(1024, 1, threads=threads)
    sc = ctx.upload(hs)
    tgt = torch.zeros((256, 256, 4), dtype=torch.uint8, device="cuda")
    acc = torch.zeros_like(tgt)
    ms = median_ms(lambda: ctx.render(sc, [0.0, 3.0, 0.0], 256, 256, vfov=60.0, bounces=4, target=tgt,
                                      accum=acc), 7)
    prim, sec = ctx.last_rays()
    out["small_render"] = {"workload": "1024-triangle soup (seed 1), 256x256, 1 spp, 4 bounces, vfov 60, "
                                       "eye (0,0,3)", "ms": round(ms, 4), "rays": prim + sec,
                           "mrays_per_s": round((prim + sec) / ms / 1e3, 2)}
    ico = rtamd.scene_icosphere(6, 2, 0.1, 4.0, threads=threads)
    isc = ctx.upload(ico)
    coh = rtamd.rays_to_device(rtamd.rays_camera_tiles([0, 3, 0], 60.0, 1024, 1024, 8, 0, 524288))
    inc = rtamd.rays_to_device(rtamd.rays_shell(524288, 2024, 3.0, 0.5))
    hits = torch.empty((524288, 4), dtype=torch.float32, device="cuda")
    t_coh = median_ms(lambda: ctx.closest_hit(isc, coh, hits), 7)
    t_inc = median_ms(lambda: ctx.closest_hit(isc, inc, hits), 7)
    out["query_medium"] = {"workload": "level-6 icosphere (81920 triangles, displacement 0.1 value noise seed 2, "
                                       "frequency 4); 524288 coherent camera rays (8x8 tiles of 1024x1024, eye "
                                       "(0,0,3), vfov 60) and 524288 incoherent shell rays (seed 2024)",
                           "coherent_ms": round(t_coh, 4), "coherent_mrays_per_s": round(524288 / t_coh / 1e3, 1),
                           "incoherent_ms": round(t_inc, 4),
                           "incoherent_mrays_per_s": round(524288 / t_inc / 1e3, 1)}
    out["doge_frame_loop"] = bench_doge_frame_loop(ctx)
    return out


DOGE_TRIANGLES = os.path.join(ROOT, "tests", "golden", "doge_triangles.npy")


def bench_doge_frame_loop(ctx, frames: int = 60, device_tree: bool = True):
    """The reference application's own workload: its doge scene (1,422
    triangles, RtScene.h:44-101) in an 800x600 window with the shader's vfov 30
    and two bounces, camera of main.cpp:37, one compute pass plus the display
    pass per frame -- the frame whose time main.cpp:406-412 prints.  With
    device_tree, the same frame on the device-built tree (RT_TREE_DEVICE) is
    timed as well."""
    import numpy as np
    import torch
    import rtamd
    tris = np.frombuffer(bytearray(np.load(DOGE_TRIANGLES).tobytes()), rtamd.TRI_DT).copy()
    mats = np.zeros(6, rtamd.MAT_DT)                                     # RtScene.h:49-62
    mats["type"] = [rtamd.RT_MAT_LAMBERTIAN] * 3 + [rtamd.RT_MAT_LIGHT, rtamd.RT_MAT_METAL, rtamd.RT_MAT_GLASS]
    mats["albedo"] = np.array([[.3, .3, .3], [.9, .1, .1], [.1, .9, .1], [2, 2, 2], [1, 1, 1], [1, 1, 1]], np.float32)
    spheres = np.zeros(1, rtamd.SPHERE_DT)                               # RtScene.h:97
    spheres["s"] = np.array([[0.6, 1.0, -1.0, 0.6]], np.float32)
    spheres["materialIndex"] = 5
    hs = rtamd.HostScene(tris, mats, rtamd.bvh_build(tris, seed=1), rtamd.scene_lights(tris, mats), spheres)
    sc = ctx.upload(hs)
    cam = [1.8, 8.6, 1.1]
    loop = rtamd.FrameLoop(ctx, sc, 800, 600, vfov=30.0, bounces=2)
    try:
        for k in range(5):
            loop.step(cam, time=0.0)
            loop.display()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for k in range(frames):
            loop.step(cam, time=0.0)
            loop.display()
        e1.record()
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / frames
        gpu_ms = e0.elapsed_time(e1) / frames
        prim, sec = ctx.last_rays()
    finally:
        loop.close()
        sc.release()
    out = {"workload": "doge scene (1422 triangles), 800x600, 1 spp per frame, 2 bounces, vfov 30, camera "
                       "(1.8, 8.6, 1.1); per frame: render, target->accumulation copy, display pass",
           "frames": frames, "ms_per_frame": round(wall_ms, 4), "gpu_ms_per_frame": round(gpu_ms, 4),
           "frames_per_s": round(1e3 / wall_ms, 1), "rays_per_frame": prim + sec,
           "mrays_per_s": round((prim + sec) / wall_ms / 1e3, 1)}
    log(f"doge frame loop: {wall_ms:.3f} ms/frame ({out['frames_per_s']} frames/s)")
    if device_tree:
        t = ctx.get_tuning()
        ctx.set_tuning(t.node_layout, t.refill, t.blocks_per_cu, t.stack_lds, rtamd.TREE_DEVICE, t.ray_order)
        try:
            out["ms_per_frame_device_tree"] = bench_doge_frame_loop(ctx, frames, device_tree=False)["ms_per_frame"]
        finally:
            ctx.set_tuning(t.node_layout, t.refill, t.blocks_per_cu, t.stack_lds, t.tree, t.ray_order)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1 << 22)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the plain-kernel measurement (no render, device build, small and medium "
                         "workloads, batched or served runs); for counter passes")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--fifo-seconds", type=float, default=10.0)
    ap.add_argument("--serve-rates", default="8000",
                    help="comma-separated arrival rates (per s) for extra 6-second runs of the batching server "
                         "(capacity beyond the 4000/s workload); empty for none")
    ap.add_argument("--layout", choices=["w4", "q16", "f32"], default="w4",
                    help="node records: 4-wide 64-byte 16-bit (default), binary 32-byte 16-bit, binary 64-byte fp32")
    ap.add_argument("--refill", type=int, default=0, help="idle lanes before a wave refetches (0 = library default)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the measurement, copy a buffer of this many MiB once (a known-size kernel for "
                         "checking the memory-traffic counters)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the counter passes (rocprofv3 --pmc runs of the query and render) and report the "
                         "committed profile's traffic instead")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--frame-loop-only", action="store_true",
                    help="only the reference's own frame (the doge scene's frame loop, 60 frames); prints its "
                         "figures as one JSON line, for kernel traces of that workload")
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args, host_threads())
        return
    if args.frame_loop_only:
        import rtamd
        print(json.dumps(bench_doge_frame_loop(rtamd.Context(0), device_tree=False)), flush=True)
        return
    if args.plain_only:
        args.no_render = args.no_batched = args.no_served = True
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rays shard by batch with no "
              "exchange; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import rtamd

    threads = host_threads()
    ctx = rtamd.Context(0)
    layout = {"w4": rtamd.LAYOUT_W4, "q16": rtamd.LAYOUT_Q16, "f32": rtamd.LAYOUT_F32}[args.layout]
    node_bytes = 32 if layout == rtamd.LAYOUT_Q16 else 64
    ctx.set_tuning(node_layout=layout, refill=args.refill)
    t0 = time.perf_counter()
    hs = rtamd.scene_clusters(args.tris, 3, threads=threads)
    log(f"scene {len(hs.triangles)} triangles built in {time.perf_counter() - t0:.1f} s")
    scene = ctx.upload(hs)
    rays_np = rtamd.rays_box(args.rays, 4, 4.0)
    rays = rtamd.rays_to_device(rays_np)
    n = len(rays_np)
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")

    # untimed counting pass: per-ray node and triangle records read
    _, nv, tt = ctx.closest_hit_stats(scene, rays)
    torch.cuda.synchronize()
    v_node = int(nv.to(torch.int64).sum().item())
    v_tri = int(tt.to(torch.int64).sum().item())
    bytes_per_launch = v_node * node_bytes + v_tri * 48 + n * (32 + 16)
    del nv, tt

    # the same counts on the canonical sample, for the binary fp32 records
    # (one record per internal node expanded) and the configured layout
    k = min(CANON_SAMPLE, n)
    sample_counts = {}
    for name, lay in (("f32", rtamd.LAYOUT_F32), (args.layout, layout)):
        ctx.set_tuning(node_layout=lay, refill=args.refill)
        _, snv, stt = ctx.closest_hit_stats(scene, rays[:k])
        sample_counts[name] = (float(snv.to(torch.float64).mean().item()), float(stt.to(torch.float64).mean().item()))
    ctx.set_tuning(node_layout=layout, refill=args.refill)

    for _ in range(args.warmup):
        ctx.closest_hit(scene, rays, hits)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        evs[k][0].record()
        ctx.closest_hit(scene, rays, hits)
        evs[k][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    log(f"plain: {ms_per_step:.3f} ms per step")
    call_ms = sum(a.elapsed_time(b) for a, b in evs) / args.steps
    # the traversal kernel alone: stage events around its launch, on its stream
    ctx.set_render_timing(True)
    stages = []
    for _ in range(max(3, min(args.steps, 10))):
        ctx.closest_hit(scene, rays, hits)
        stages.append(ctx.last_query_stats())
    ctx.set_render_timing(False)
    kernel_ms = sum(st.trace_ms for st in stages) / len(stages)
    order_ms = sum(st.order_ms for st in stages) / len(stages)
    sorted_rays = bool(stages[0].sorted)
    value = n / (ms_per_step / 1e3) / 1e6
    record_gbs = bytes_per_launch / (kernel_ms / 1e3) / 1e9
    copy_gbs = copy_bandwidth()

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"closest-hit query, clustered soup {len(hs.triangles)} triangles in 4096 Gaussian "
                               f"clusters, {n} incoherent rays (origins U[-4,4]^3, uniform directions)",
                   "triangles": len(hs.triangles), "rays": n,
                   "traversal": f"ordered, persistent, {args.layout} {node_bytes}-byte node records",
                   "trace_order": "origin Morton order (sorted on the device inside the timed call)"
                   if sorted_rays else "index order"},
        "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "traffic": None, "kernel": "k_trace_packed", "kernel_ms": round(kernel_ms, 4),
                     "call_ms": round(call_ms, 4), "order_ms": round(order_ms, 4),
                     # every node and triangle record fetch priced at its record size (hits in
                     # the caches included), per launch and over the kernel's time
                     "record_fetch_bytes_per_launch": bytes_per_launch, "record_fetch_gbs": round(record_gbs, 1),
                     "record_fetch_frac": round(record_gbs / HBM_PEAK_GBS, 4),
                     "stream_copy_gbs": round(copy_gbs, 1)},
        "visits": {"node_records_per_ray": round(v_node / n, 2), "tri_records_per_ray": round(v_tri / n, 3)},
    }

    if not args.no_render:
        tgt = torch.zeros((1080, 1920, 4), dtype=torch.uint8, device="cuda")
        acc = torch.zeros_like(tgt)
        ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
        e1.record()
        torch.cuda.synchronize()
        rms = e0.elapsed_time(e1)
        prim, sec = ctx.last_rays()
        ctx.set_render_timing(True)
        ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
        st = ctx.last_render_stats()
        ctx.set_render_timing(False)
        # the render's record fetches, from one untimed counted render
        ctx.set_render_counting(True)
        ctx.render_frames(scene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
        r_nv, r_nt = ctx.last_render_counts()
        ctx.set_render_counting(False)
        out["render"] = {"workload": "1920x1080, 4 spp (currentSample 0..3), 8 bounces, vfov 60, eye (0,0,10)",
                         "ms": round(rms, 3), "rays": prim + sec, "secondary_rays": sec,
                         "mrays_per_s": round((prim + sec) / (rms / 1e3) / 1e6, 2),
                         "trace_primary_ms": round(st.trace_primary_ms, 3),
                         "trace_secondary_ms": round(st.trace_secondary_ms, 3),
                         "shade_ms": round(st.shade_ms, 3),
                         "primary_trace_mrays_per_s": round(prim / st.trace_primary_ms / 1e3, 1),
                         "secondary_trace_mrays_per_s": round(sec / st.trace_secondary_ms / 1e3, 1)}
        r_bytes = r_nv * node_bytes + r_nt * 48 + (prim + sec) * (32 + 16)
        r_trace_s = (st.trace_primary_ms + st.trace_secondary_ms) / 1e3
        out["render"].update({"node_records_per_ray": round(r_nv / (prim + sec), 2),
                              "tri_records_per_ray": round(r_nt / (prim + sec), 3),
                              "record_fetch_bytes": int(r_bytes),
                              "record_fetch_gbs": round(r_bytes / r_trace_s / 1e9, 1),
                              "record_fetch_frac": round(r_bytes / r_trace_s / 1e9 / HBM_PEAK_GBS, 4)})
        log(f"render: {rms:.1f} ms, secondary traversal {out['render']['secondary_trace_mrays_per_s']} Mrays/s")
        if not args.plain_only:
            # the same render on the device-built tree (RT_TREE_DEVICE): its
            # time, and how far its image is from the given tree's (only the
            # choice between exactly equal hits can differ)
            given = tgt.cpu().numpy().astype("float64")
            t = ctx.get_tuning()
            ctx.set_tuning(t.node_layout, t.refill, t.blocks_per_cu, t.stack_lds, rtamd.TREE_DEVICE, t.ray_order)
            try:
                dscene = ctx.upload(hs)
            finally:
                ctx.set_tuning(t.node_layout, t.refill, t.blocks_per_cu, t.stack_lds, t.tree, t.ray_order)
            tgt.zero_()
            acc.zero_()
            ctx.render_frames(dscene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
            torch.cuda.synchronize()
            e0.record()
            ctx.render_frames(dscene, [0.0, 10.0, 0.0], 1920, 1080, 4, vfov=60.0, bounces=8, target=tgt, accum=acc)
            e1.record()
            torch.cuda.synchronize()
            d = (tgt.cpu().numpy().astype("float64") - given) / 255.0
            out["render"]["device_tree_ms"] = round(e0.elapsed_time(e1), 3)
            out["render"]["device_tree_rmse"] = float("%.3g" % float((d * d).mean(axis=(0, 1)).max() ** 0.5))
            log(f"render on the device-built tree: {out['render']['device_tree_ms']} ms, "
                f"rmse {out['render']['device_tree_rmse']} against the given tree's image")
            dscene.release()

    # counters of the timed kernels (and the render's), in child runs under rocprofv3
    agg = None if args.no_pmc else run_counter_passes(args)
    num_cus = torch.cuda.get_device_properties(0).multi_processor_count
    rf = out["roofline"]
    q = counter_summary(agg["query"], num_cus) if agg and "query" in agg else {}
    cal = counter_summary(agg["copy"], num_cus) if agg and "copy" in agg else {}
    cal_read = cal_write = 1.0
    if cal and cal["dram_read_bytes"] > 0 and cal["dram_write_bytes"] > 0:
        # a copy of known size reads and writes CALIBRATION_MIB MiB: scale the counts by what it shows
        known = float(CALIBRATION_MIB << 20)
        cal_read, cal_write = known / cal["dram_read_bytes"], known / cal["dram_write_bytes"]
        rf["traffic_calibration"] = {"copy_mib": CALIBRATION_MIB, "read_factor": round(cal_read, 4),
                                     "write_factor": round(cal_write, 4)}
    if q:
        q["dram_bytes"] = q["dram_read_bytes"] * cal_read + q["dram_write_bytes"] * cal_write
    if q:
        traffic = q["dram_bytes"]
        rf["traffic_source"] = ("measured in this run: rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B_sum, "
                                "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum, TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum (x 32 bytes), "
                                f"mean of {q['dispatches']} launches")
    else:
        traffic, current = read_traffic()
        rf["traffic_source"] = (f"{os.path.relpath(TRAFFIC_FILE, ROOT)} (committed profile; "
                                + ("taken on these kernel sources)" if current else "STALE: the kernel sources "
                                   "changed since it was taken)")) if traffic else "none"
    if traffic:
        achieved = traffic / (kernel_ms / 1e3) / 1e9
        rf.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                   "traffic": int(traffic), "traffic_vs_record_fetch": round(traffic / bytes_per_launch, 4)})
    if q:
        rf["l2_hit"] = round(q["l2_hit"], 4) if q["l2_hit"] is not None else None
        if "profiled_kernel_ms" in q:
            rf["profiled_kernel_ms"] = round(q["profiled_kernel_ms"], 4)
        # what binds the loop: VALU issue, not bandwidth
        rf["valu"] = {"busy": round(q["valu_busy"], 4) if q["valu_busy"] else None,
                      "lanes_active": round(q["valu_lanes_active"], 4) if q["valu_lanes_active"] else None,
                      "wave_insts_per_ray": round(q["valu_insts"] / n, 1),
                      "salu_insts_per_ray": round(q["salu_insts"] / n, 1),
                      "wave_insts_per_record": round(q["valu_insts"] / max(1, v_node + v_tri), 2)}
        rf["binding_resource"] = "valu issue" if (q["valu_busy"] or 0) > 0.8 else "memory latency"
    if agg and "render" in out:
        r = out["render"]
        rays_r = r["rays"]
        tr = counter_summary(agg.get("render_trace", {}), num_cus)
        sh = counter_summary(agg.get("shade", {}), num_cus)
        for cs in (tr, sh):
            if cs:
                cs["dram_bytes"] = cs["dram_read_bytes"] * cal_read + cs["dram_write_bytes"] * cal_write
        rc = {}
        for name, cs, ms in (("trace", tr, r["trace_primary_ms"] + r["trace_secondary_ms"]), ("shade", sh, r["shade_ms"])):
            if not cs:
                continue
            tot = lambda key: cs[key] * cs["dispatches"]
            rc[name] = {"dispatches": cs["dispatches"], "dram_bytes": int(tot("dram_bytes")),
                        "dram_bytes_per_ray": round(tot("dram_bytes") / rays_r, 1),
                        "dram_gbs": round(tot("dram_bytes") / (ms / 1e3) / 1e9, 1),
                        "dram_frac": round(tot("dram_bytes") / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "l2_hit": round(cs["l2_hit"], 4) if cs["l2_hit"] is not None else None,
                        "valu_busy": round(cs["valu_busy"], 4) if cs["valu_busy"] else None,
                        "valu_lanes_active": round(cs["valu_lanes_active"], 4) if cs["valu_lanes_active"] else None,
                        "valu_wave_insts_per_ray": round(tot("valu_insts") / rays_r, 1)}
        if rc:
            r["counters"] = dict(rc, source="measured in this run (rocprofv3 --pmc, one render), per ray of the "
                                            "render's primary + secondary rays, over the stage times above")

    if not args.plain_only:
        out["device_build"] = bench_device_build(ctx, hs, rays, threads, layout, node_bytes)
        out.update(bench_small_and_medium(ctx, threads))

    if not args.no_batched:
        out["batched"] = bench_batched(ctx, threads, 5)
    if not args.no_served:
        rates = [int(x) for x in args.serve_rates.split(",") if x.strip()]
        out["served"] = bench_served(ctx, threads, args.serve_seconds, args.fifo_seconds, rates)

    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = src.numel() * 4
        del src, dst

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(hs, rays_np, args.cpu_seconds, threads)
        canon = out["cpu_baseline"]["canonical"]
        f32v, f32t = sample_counts["f32"]
        lv, lt = sample_counts[args.layout]
        # wasted-work guard: the kernel's binary traversal against the
        # canonical ordered one (visits), and the bytes per ray that
        # `achieved` counts against the canonical visits priced at the most
        # compact binary record (32 B) + 48 B triangles + 48 B ray and hit
        canon_bytes = canon["node_visits_per_ray"] * 32 + canon["tri_tests_per_ray"] * 48 + 48
        layout_bytes = lv * node_bytes + lt * 48 + 48
        out["roofline"]["visit_guard"] = {
            "rays": canon["rays"],
            "f32_node_visits_per_ray": round(f32v, 3), "f32_tri_tests_per_ray": round(f32t, 3),
            "f32_visits_vs_canonical": round(f32v / canon["node_visits_per_ray"], 4),
            f"{args.layout}_records_per_ray": round(lv, 3), f"{args.layout}_tri_tests_per_ray": round(lt, 3),
            f"{args.layout}_tri_tests_vs_canonical": round(lt / canon["tri_tests_per_ray"], 4),
            "canonical_bytes_per_ray": round(canon_bytes, 1), f"{args.layout}_bytes_per_ray": round(layout_bytes, 1),
            "bytes_vs_canonical": round(layout_bytes / canon_bytes, 4),
            "credited": bool(f32v <= 1.05 * canon["node_visits_per_ray"] and layout_bytes <= 1.05 * canon_bytes)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
