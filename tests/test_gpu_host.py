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
