"""bench.py's host side on the CPU: its seeded generator reproduces the golden
fixtures' operands (so the bench's products are checked against GMP digests), the
kernel names it forms match the committed rocprofv3 tables it reads the HBM
traffic and VALU counts from, and a request for more than one GPU is refused."""
import json
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.fixture(scope="module")
def cases():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


@pytest.mark.parametrize("name", ["config-n65536-w2048-00", "config-n65536-w2048-15", "config-uneven-00",
                                  "config-uneven-63"])
def test_generator_reproduces_golden_operands(cases, name):
    c = cases[name]
    q, a, b = bench.make_point(c["seed"], c["N"], c["W"])
    assert "%x" % sum(int(w) << (32 * i) for i, w in enumerate(q)) == c["q_hex"]
    assert bench.digest(a) == c["a_sha"] and bench.digest(b) == c["b_sha"]


@pytest.mark.parametrize("key,name,_ms", bench.REFERENCE_POINTS)
def test_reference_points_reproduce_golden_operands(cases, key, name, _ms):
    import numpy as np
    c = cases[name]
    qi = int(c["q_hex"], 16)
    q = np.array([(qi >> (32 * l)) & 0xFFFFFFFF for l in range(c["L"])], dtype=np.uint32)
    _, a, b = bench.make_operands(c["seed"], c["N"], c["W"], q)
    assert bench.digest(a) == c["a_sha"] and bench.digest(b) == c["b_sha"]


def test_golden_digests_cover_the_batch():
    gold = bench.golden_digests()
    assert all(bench.SEED0 + i in gold for i in range(bench.BATCH))


@pytest.mark.parametrize("fused,rows", [(1, "k_ntt_rows_fused_mid_batch<true,1>"),
                                        (0, "k_ntt_rows_fused_reg<8,true,1>")])
def test_kernel_names_match_the_committed_profiles(fused, rows):
    plan = SimpleNamespace(twiddles=1, K=134, fused_cols=fused)
    names = bench.phase_kernels(plan)
    assert names["ntt_rows_fused"] == rows
    assert ("ntt_cols_fwd_a" in names) == (fused == 0)
    for phase, kernel in names.items():
        assert bench.measured_traffic(kernel) is not None, (phase, kernel)
        if phase.startswith("ntt"):
            assert bench.valu_issue(kernel, 0.05, 256) is not None, (phase, kernel)


def _counters(path, rows):
    with open(path, "w") as f:
        f.write("Dispatch_Id,Grid_Size,Kernel_Name,Counter_Name,Counter_Value\n")
        for d, grid, name, counter, value in rows:
            f.write('%d,%d,"%s",%s,%r\n' % (d, grid, name, counter, value))


def test_traffic_table_scales_by_the_yardsticks(tmp_path):
    """The read and write factors come from the yardsticks (here 2 and 1), a
    decompose launch over both operands is keyed apart from one over one."""
    n = bench.N
    fetch = [(1, 1, "k_copy16(x)", "FETCH_SIZE", 131072.0), (2, 1, "k_sum16(x)", "FETCH_SIZE", 131072.0),
             (3, n, "void pm::k_decompose_cols_batch<4, 1>(P)", "FETCH_SIZE", 100.0),
             (4, 2 * n, "void pm::k_decompose_cols_batch<4, 1>(P)", "FETCH_SIZE", 300.0),
             (5, 1, "void pm::k_crt_mfma<9>(P)", "FETCH_SIZE", 50.0)]
    write = [(1, 1, "k_copy16(x)", "WRITE_SIZE", 262144.0), (2, 1, "k_sum16(x)", "WRITE_SIZE", 1.0),
             (3, n, "void pm::k_decompose_cols_batch<4, 1>(P)", "WRITE_SIZE", 10.0),
             (4, 2 * n, "void pm::k_decompose_cols_batch<4, 1>(P)", "WRITE_SIZE", 20.0),
             (5, 1, "void pm::k_crt_mfma<9>(P)", "WRITE_SIZE", 5.0)]
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    _counters(fp, fetch)
    _counters(wp, write)
    out = bench.traffic_table(str(fp), str(wp), lambda *a: None)
    assert out["k_decompose_cols_batch<4,1>@1"]["bytes"] == (2 * 100 + 10) * 1024
    assert out["k_decompose_cols_batch<4,1>@2"]["bytes"] == (2 * 300 + 20) * 1024
    assert out["k_decompose_cols_batch<4,1>@2"]["algorithmic"] == 2 * out["k_decompose_cols_batch<4,1>@1"]["algorithmic"]
    assert out["k_crt_mfma<9>"]["bytes"] == (2 * 50 + 5) * 1024


def test_sq_table_per_wave(tmp_path):
    rows = [(1, 1, "void pm::k_crt_mfma<9>(P)", "SQ_WAVES", 2048.0),
            (1, 1, "void pm::k_crt_mfma<9>(P)", "SQ_INSTS_VALU", 2048.0 * 2522),
            (2, 1, "void pm::k_crt_mfma<9>(P)", "SQ_WAVES", 2048.0),
            (2, 1, "void pm::k_crt_mfma<9>(P)", "SQ_INSTS_VALU", 2048.0 * 2524)]
    path = tmp_path / "sq.csv"
    _counters(path, rows)
    out = bench.sq_table([str(path)], lambda *a: None)
    assert out["k_crt_mfma<9>"]["waves"] == 2048 and out["k_crt_mfma<9>"]["SQ_INSTS_VALU"] == 2523.0


def test_roofline_helpers_read_the_profiles():
    rows = "k_ntt_rows_fused_reg<8,true,1>"
    traffic = bench.measured_traffic(rows)
    alg = 12 * 134 * 65536
    assert alg <= traffic <= 1.2 * alg
    v = bench.valu_issue(rows, 0.044, 256)
    assert v["waves"] == 8576 and 2000 < v["instructions_per_wave"] < 3000
    # the assembly's class counts add up to the counter's VALU instructions per wave
    for kernel, mix in bench.VALU_MIX.items():
        w = bench.valu_issue(kernel, 0.02, 256)
        assert sum(mix.values()) == round(w["instructions_per_wave"]), kernel
    # 1584 plain ops at 2.92 cycles, 448 high and 321 low multiplies at 4.64 and 5.04, one 64-bit multiply-add
    cycles = 1584 * 2.92 + 448 * 4.64 + 321 * 5.04 + 5.9
    expect = cycles * 8576 / 1024 / 2.4e9 * 1e3
    assert abs(v["issue_ms"] - expect) < 1e-4 and abs(v["frac"] - expect / 0.044) < 1e-3


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def _trace(path, rows):
    with open(path, "w") as f:
        f.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for name, s, e in rows:
            f.write('"%s",%d,%d\n' % (name, s, e))


def test_trace_shares_splits_overlap_evenly(tmp_path):
    """Three replays of two overlapping kernels and a runtime copy: only the last
    replay counts, the copy is left out, and time run by both is halved."""
    rows = []
    for r in range(3):
        t = 10000 * r
        rows += [("void pm::(anonymous namespace)::k_a<8, true>(pm::DevPlan)", t, t + 1000),
                 ("void pm::(anonymous namespace)::k_b<9>(pm::DevPlan)", t + 500, t + 2000),
                 ("__amd_rocclr_copyBuffer", t + 3000, t + 4000)]
    path = tmp_path / "trace.csv"
    _trace(path, rows)
    out = bench.trace_shares(str(path), 3, lambda *a: None)
    assert out["dispatches_per_replay"] == 2 and out["idle_us"] == 0
    assert out["replay_us"] == 2.0
    assert out["kernels"]["k_a"]["us"] == 0.75 and out["kernels"]["k_b"]["us"] == 1.25
    assert abs(sum(v["share"] for v in out["kernels"].values()) - 1) < 1e-9
    with pytest.raises(SystemExit):
        bench.trace_shares(str(path), 4, lambda *a: None)


def test_trace_stats_median_and_trim(tmp_path):
    rows = [("void pm::k_c<1>(int)", 0, d) for d in [1000] * 18 + [9000, 100]]
    path = tmp_path / "trace.csv"
    _trace(path, rows)
    out = bench.trace_stats(str(path), lambda *a: None)["k_c<1>"]
    assert out["launches"] == 20 and out["median_us"] == 1.0 and out["trimmed_mean_us"] == 1.0
    assert out["min_us"] == 0.1


def test_clock_table_divides_by_xcds_and_duration(tmp_path):
    """GRBM_COUNT over 8 XCDs and the dispatch's duration; short dispatches and
    other counters are left out."""
    path = tmp_path / "clk.csv"
    with open(path, "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp\n")
        f.write('1,"void pm::k_e<2>(int)",GRBM_COUNT,%r,0,500000\n' % (8 * 2.0 * 500000))
        f.write('2,"void pm::k_e<2>(int)",GRBM_COUNT,%r,0,1000000\n' % (8 * 1.9 * 1000000))
        f.write('2,"void pm::k_e<2>(int)",SQ_WAVES,64.0,0,1000000\n')
        f.write('3,"void pm::k_e<2>(int)",GRBM_COUNT,%r,0,100000\n' % (8 * 9.0 * 100000))
    out = bench.clock_table(str(path), lambda *a: None)
    assert out["k_e<2>"]["dispatches"] == 2 and abs(out["k_e<2>"]["ghz"] - 1.95) < 1e-9
    assert out["k_e<2>"]["mean_us"] == 750.0
