"""The bench's counter bookkeeping on synthetic rocprofv3 rows (no GPU): the
kernel classes, the per-dispatch means, the copy that calibrates the byte
counts (its largest dispatch only), and the derived figures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

QUERY = "void rtamd::(anonymous namespace)::k_trace_packed<256, 36, 4, false, 0>(rtamd::TraceScene, rtamd::TraceIo, int*)"
COUNTING = "void rtamd::(anonymous namespace)::k_trace_packed<256, 36, 4, true, 0>(rtamd::TraceScene, rtamd::TraceIo, int*)"
RENDER = "void rtamd::(anonymous namespace)::k_trace_render<256, 36, 4, 0>(rtamd::TraceScene, rtamd::TraceIo, int*)"


def rows(kernel, dispatch, grid, counters, t0=0, t1=1000):
    return [{"Kernel_Name": kernel, "Dispatch_Id": str(dispatch), "Grid_Size": str(grid), "Counter_Name": k,
             "Counter_Value": str(v), "Start_Timestamp": str(t0), "End_Timestamp": str(t1)} for k, v in counters.items()]


def test_counter_passes_and_calibration():
    p0 = {"TCC_EA0_RDREQ_DRAM_32B_sum": 1000, "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum": 10,
          "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum": 2, "TCC_HIT_sum": 300}
    p1 = {"SQ_WAVES": 64, "SQ_INSTS_VALU": 4000, "SQ_INSTS_SALU": 100, "SQ_BUSY_CYCLES": 32000,
          "SQ_ACTIVE_INST_VALU": 4000, "SQ_THREAD_CYCLES_VALU": 128000, "TCC_MISS_sum": 100}
    agg = {}
    r0 = rows(QUERY, 1, 4096, p0, 0, 2000) + rows(QUERY, 2, 4096, {k: 3 * v for k, v in p0.items()}, 0, 4000)
    r0 += rows(COUNTING, 3, 4096, {k: 99999 for k in p0})                         # not the timed kernel
    r0 += rows(RENDER, 4, 4096, p0)
    # a small transfer and the 1 GiB calibration copy
    r0 += rows("__amd_rocclr_copyBuffer", 5, 512, {k: 7 for k in p0})
    r0 += rows("__amd_rocclr_copyBuffer", 6, 131072, {"TCC_EA0_RDREQ_DRAM_32B_sum": (1 << 30) // 32,
                                                    "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum": (1 << 30) // 64,
                                                    "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum": 0, "TCC_HIT_sum": 1})
    bench.add_counter_rows(agg, r0, 0)
    r1 = rows(QUERY, 1, 4096, p1) + rows(QUERY, 2, 4096, p1) + rows(RENDER, 4, 4096, p1)
    r1 += rows("__amd_rocclr_copyBuffer", 6, 131072, p1) + rows("__amd_rocclr_copyBuffer", 5, 512, p1)
    bench.add_counter_rows(agg, r1, 1)
    assert set(agg) == {"query", "render_trace", "copy"}
    q = bench.counter_summary(agg["query"], 256)
    assert q["dispatches"] == 2
    assert q["dram_read_bytes"] == pytest.approx(32 * 2000)          # mean of 1000 and 3000
    assert q["dram_write_bytes"] == pytest.approx(32 * 24)
    assert q["l2_hit"] == pytest.approx(600 / 700)
    assert q["valu_busy"] == pytest.approx(4000 * 4 / (1024 * 32000 / 32))
    assert q["valu_lanes_active"] == pytest.approx(0.5)
    assert q["profiled_kernel_ms"] == pytest.approx(3000 / 1e6)
    c = bench.counter_summary(agg["copy"], 256)
    assert c["dispatches"] == 1                                      # the small transfer is left out
    assert c["dram_read_bytes"] == 1 << 30 and c["dram_write_bytes"] == 1 << 29


def test_traffic_file_marks_stale_sources(tmp_path, monkeypatch):
    f = tmp_path / "traffic.txt"
    f.write_text("# x\nbytes_per_launch=123\nkernel_sources_sha=0000\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.read_traffic() == (123.0, False)
    f.write_text(f"bytes_per_launch=5\nkernel_sources_sha={bench.kernel_sources_sha()}\n")
    assert bench.read_traffic() == (5.0, True)
