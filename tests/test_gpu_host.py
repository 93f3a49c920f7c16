"""GPU: products on host memory through every host-transfer mode, the timed
mode choice, and the FHERMA C++ boundary through the point-directory driver."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import pointgen

polymult = pytest.importorskip("polymult")

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOLUTION = os.path.join(ROOT, "fherma-polymult_amd", "solution")
MODES = [("staged", 0), ("mapped", 1), ("direct", 2)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [m for _, m in MODES], ids=[n for n, _ in MODES])
@pytest.mark.parametrize("name", ["matrix-n2-w33-random", "matrix-n64-w2048-random", "config-n4096-w256-s1",
                                  "config-n32768-w868-s1", "config-n65536-w2048-01", "matrix-n131072-w512-prime"])
@pytest.mark.parametrize("form", [0, 1], ids=["plain", "mfma"])
def test_host_modes_match_gmp(golden, name, mode, form):
    case = golden[name]
    q, a, b = pointgen.operands(case)
    say("host", name, "mode", mode, "form", form)
    with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
        plan.set_host_mode(mode)
        for _ in range(2):   # the second call reuses the pipeline
            c = plan.multiply(a, b)
            assert oracle.digest(c) == case["c_sha"], (name, mode)


@pytest.mark.parametrize("chunks", [1, 3, 16])
@pytest.mark.parametrize("mode", [0, 1], ids=["staged", "mapped"])
def test_host_chunk_counts_match_gmp(golden, chunks, mode, monkeypatch):
    """PM_HOST_CHUNKS fixes the coefficient chunks per operand; the timed choice
    then keeps that count and picks only the mode."""
    case = golden["config-n32768-w868-s1"]
    q, a, b = pointgen.operands(case)
    monkeypatch.setenv("PM_HOST_CHUNKS", str(chunks))
    say("host chunks", chunks, "mode", mode)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        plan.set_host_mode(mode)
        assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]
        tuned, ms = plan.tune_host()
        assert tuned in (0, 1, 2) and ms > 0
        assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]


def test_timed_mode_choice(golden):
    case = golden["config-n32768-w868-s2"]
    q, a, b = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        mode, ms = plan.tune_host()
        say("tuned mode", mode, "%.3f ms" % ms)
        assert mode in (0, 1, 2) and ms > 0
        assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]


def test_timed_kernel_choice(golden, monkeypatch):
    """pm_plan_tune_kernels times the other twiddle source and, at N = 65536,
    the other column-pass form; the settings it keeps give the GMP product, and
    FHERMA_TWIDDLES / PM_NTT_FORM pin their dimension."""
    for name in ("config-n65536-w2048-01", "config-n32768-w868-s1"):
        case = golden[name]
        q, a, b = pointgen.operands(case)
        with polymult.Plan(case["N"], case["W"], q) as plan:
            ms = plan.tune_kernels()
            say(name, "kept twiddles", plan.twiddles, "ntt form", plan.ntt_form, "%.3f ms" % ms)
            assert ms > 0
            assert plan.twiddles in (polymult.TWIDDLES_TABLE, polymult.TWIDDLES_GEN)
            if case["N"] != 65536:
                assert plan.ntt_form == polymult.NTT_VALU
            assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]
    case = golden["config-n65536-w2048-01"]
    q, a, b = pointgen.operands(case)
    monkeypatch.setenv("FHERMA_TWIDDLES", "table")
    monkeypatch.setenv("PM_NTT_FORM", "mfma")
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.tune_kernels() > 0
        assert plan.twiddles == polymult.TWIDDLES_TABLE and plan.ntt_form == polymult.NTT_MFMA
        assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]


def write_point(root, case, patterns=("random",)):
    q, a, b = pointgen.operands(case)
    os.makedirs(os.path.join(root, "point"), exist_ok=True)
    q.astype("<u4").tofile(os.path.join(root, "point", "q.bin"))
    with open(os.path.join(root, "manifest.json"), "w") as f:
        json.dump({"N": case["N"], "W": case["W"], "L": case["L"], "cases": 2}, f)
    for i in range(2):
        d = os.path.join(root, "cases", "%06d" % i)
        os.makedirs(d, exist_ok=True)
        a.astype("<u4").tofile(os.path.join(d, "a.bin"))
        b.astype("<u4").tofile(os.path.join(d, "b.bin"))


@pytest.mark.parametrize("name", ["config-n4096-w256-s3", "matrix-n32768-w868-prime", "config-n65536-w1747-s1"])
def test_point_directory_driver(golden, name, tmp_path):
    """fherma_init / fherma_run / fherma_free through the harness protocol."""
    case = golden[name]
    write_point(str(tmp_path), case)
    say("driver", name)
    r = subprocess.run([SOLUTION, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.load(open(os.path.join(tmp_path, "out", "results.json")))
    assert [c["status"] for c in res["cases"]] == ["ok", "ok"]
    for i in range(2):
        c = np.fromfile(os.path.join(tmp_path, "out", "%06d" % i, "c.bin"), dtype="<u4")
        assert oracle.digest(c) == case["c_sha"]


def test_driver_refuses_bad_operand_size(golden, tmp_path):
    case = golden["config-n4096-w256-s1"]
    write_point(str(tmp_path), case)
    np.zeros(7, dtype="<u4").tofile(os.path.join(tmp_path, "cases", "000001", "a.bin"))
    r = subprocess.run([SOLUTION, str(tmp_path)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, PM_AUTOTUNE="0"))
    assert r.returncode == 0
    res = json.load(open(os.path.join(tmp_path, "out", "results.json")))
    assert res["cases"][0]["status"] == "ok" and res["cases"][1]["status"] == "crashed"


@pytest.mark.parametrize("out_mode", [m for _, m in MODES], ids=[n for n, _ in MODES])
@pytest.mark.parametrize("in_mode", [m for _, m in MODES], ids=[n for n, _ in MODES])
def test_mixed_host_modes_match_gmp(golden, in_mode, out_mode):
    """The input and output sides in different modes (the reference's FHERMA_IN
    and FHERMA_OUT), every pair."""
    case = golden["config-n32768-w868-s1"]
    q, a, b = pointgen.operands(case)
    say("host modes in", in_mode, "out", out_mode)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        plan.set_host_modes(in_mode, out_mode)
        assert plan.host_modes == (in_mode, out_mode)
        for _ in range(2):
            assert oracle.digest(plan.multiply(a, b)) == case["c_sha"]


def run_driver(tmp_path, case, env):
    write_point(str(tmp_path), case)
    e = dict(os.environ)
    for k in [k for k in e if k.startswith("FHERMA_") or k.startswith("PM_")]:
        del e[k]
    e.update(env)
    return subprocess.run([SOLUTION, str(tmp_path)], capture_output=True, text=True, timeout=120, env=e)


def answers_ok(tmp_path, case):
    res = json.load(open(os.path.join(tmp_path, "out", "results.json")))
    assert [c["status"] for c in res["cases"]] == ["ok", "ok"]
    for i in range(2):
        c = np.fromfile(os.path.join(tmp_path, "out", "%06d" % i, "c.bin"), dtype="<u4")
        assert oracle.digest(c) == case["c_sha"]


@pytest.mark.parametrize("env", [
    {"FHERMA_IN": "staged", "FHERMA_OUT": "staged"},
    {"FHERMA_IN": "mapped", "FHERMA_OUT": "staged"},
    {"FHERMA_IN": "register", "FHERMA_OUT": "mapped"},
    {"FHERMA_IN": "hmm", "FHERMA_OUT": "stream"},
    {"FHERMA_IN": "staged", "FHERMA_CHUNKS": "3", "FHERMA_OUT_CHUNKS": "5"},
    {"FHERMA_IN": "mapped", "FHERMA_OUT": "mapped", "FHERMA_A_CHUNKS": "1,3", "FHERMA_B_CHUNKS": "2,1,1",
     "FHERMA_C_CHUNKS": "1,2,1"},
    {"FHERMA_AUTOTUNE": "0", "FHERMA_THREADS": "3", "FHERMA_SPIN_MS": "5"},
], ids=["staged", "mapped-staged", "register-mapped", "hmm-stream", "chunks", "weights", "no-autotune"])
def test_driver_honours_the_reference_knobs(golden, tmp_path, env):
    """fherma_init reads the reference's host knobs (solve.cu:13-49): the copy
    modes (hmm and stream fall back to staged, register is a DMA from the
    caller's memory), chunk counts and weights, threads; every answer vs GMP."""
    case = golden["matrix-n32768-w868-prime"]
    say("driver env", env)
    r = run_driver(tmp_path, case, env)
    assert r.returncode == 0, r.stderr
    answers_ok(tmp_path, case)


def test_driver_refuses_unknown_copy_mode(golden, tmp_path):
    r = run_driver(tmp_path, golden["config-n4096-w256-s1"], {"FHERMA_IN": "pinned"})
    assert r.returncode != 0 and "FHERMA_IN must be staged, register, mapped, hmm or stream" in r.stderr


def test_driver_autotune_switch_and_profile(golden, tmp_path):
    """FHERMA_AUTOTUNE_LOG lists the timed candidates (host modes, then the
    twiddle sources) unless FHERMA_AUTOTUNE=0 skips the choice; FHERMA_PROFILE prints the phase times of every call."""
    case = golden["config-n4096-w256-s1"]
    r = run_driver(tmp_path, case, {"FHERMA_AUTOTUNE_LOG": "1", "FHERMA_PROFILE": "1"})
    assert r.returncode == 0, r.stderr
    assert r.stderr.count("host mode in") >= 8 and r.stderr.count("gpu us: in+decompose") == 2, r.stderr
    assert r.stderr.count("device twiddles") >= 2, r.stderr
    answers_ok(tmp_path, case)
    r = run_driver(tmp_path, case, {"FHERMA_AUTOTUNE_LOG": "1", "FHERMA_AUTOTUNE": "0"})
    assert r.returncode == 0 and "host mode in" not in r.stderr and "device twiddles" not in r.stderr, r.stderr
    answers_ok(tmp_path, case)


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_host_batch_matches_gmp(golden, depth):
    """pm_batch_multiply_host: products of numpy operands, `depth` in flight,
    mixed shapes and host modes, one plan used twice (its products serialize);
    every answer against its GMP digest."""
    names = ["config-n65536-w2048-00", "config-n65536-w2048-01", "config-n32768-w868-s1",
             "config-n4096-w256-s1", "config-n65536-w1747-s1"]
    cases = [golden[n] for n in names]
    ops = [pointgen.operands(c) for c in cases]
    plans = [polymult.Plan(c["N"], c["W"], q) for c, (q, _, _) in zip(cases, ops)]
    plans[1].set_host_modes(polymult.HOST_MAPPED, polymult.HOST_STAGED)
    plans[2].set_host_modes(polymult.HOST_DIRECT, polymult.HOST_MAPPED)
    # the first plan twice: its second product squares case 0's a
    order = [0, 1, 2, 3, 4, 0]
    use_plans = [plans[i] for i in order]
    A = [ops[i][1] for i in order[:5]] + [ops[0][1]]
    B = [ops[i][2] for i in order[:5]] + [ops[0][1]]
    say("host batch depth", depth)
    out = polymult.batch_multiply_host(use_plans, A, B, depth=depth)
    for k, i in enumerate(order[:5]):
        assert oracle.digest(out[k]) == cases[i]["c_sha"], (depth, cases[i]["name"])
    want = oracle.Plan(cases[0]["N"], ops[0][0]).multiply(ops[0][1], ops[0][1])
    assert np.array_equal(out[5], want)
    for p in plans:
        p.close()
