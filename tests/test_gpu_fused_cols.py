"""Products whose first forward column stages run inside the decompose kernel
and the next ones inside the row pass (fused column stages, N = 65536 and
32768), on the
GPU: one product at a time on the device path, the allocation-free path,
batched launches over products of different K and L, and the captured batch,
each product against its GMP digest and against the same plan with the fused
stages off.
"""
import sys

import numpy as np
import pytest

import oracle
import pointgen

polymult = pytest.importorskip("polymult")

pytestmark = pytest.mark.gpu

# N = 65536 with L = 64, 55, 63, 53, 51, 44, 36, 23, 22, 11 and 6 (every K-step
# count of the decompose, L not a multiple of 4 included), the reference's
# prime q, all-(q - 1) operands and the adversarial patterns.
CASES = ["config-n65536-w2048-00", "config-n65536-w1747-s1", "matrix-n65536-w1747-prime",
         "matrix-n65536-w1747-prime-allqm1", "pattern-all_q_minus_1-n65536-w2048", "pattern-wrap-n65536-w2048",
         "pattern-alternating-n65536-w2048", "pattern-identity-n65536-w2048", "config-uneven-37",
         "config-uneven-59", "config-uneven-35", "config-uneven-28", "config-uneven-24", "config-uneven-45",
         "config-uneven-40", "config-uneven-11", "config-uneven-15"]
# N = 32768 (three outer stages): the reference's second point (868-bit prime q),
# random q, all-(q - 1) operands, the patterns and the uneven workload's products.
CASES_32K = ["config-n32768-w868-s1", "matrix-n32768-w868-prime", "matrix-n32768-w868-prime-allqm1",
             "pattern-all_q_minus_1-n32768-w868", "pattern-wrap-n32768-w868", "pattern-alternating-n32768-w868",
             "config-uneven-02", "config-uneven-03", "config-uneven-19", "config-uneven-23", "config-uneven-46",
             "config-uneven-47", "config-uneven-54", "config-uneven-55", "config-uneven-56"]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", CASES + CASES_32K)
def test_fused_cols_device_product_matches_gmp(golden, name):
    case = golden[name]
    q, a, b = pointgen.operands(case)
    N, L = case["N"], case["L"]
    say("fused cols", name, "L", L)
    with polymult.Plan(N, case["W"], q, form=polymult.FORM_MFMA) as plan:
        assert plan.fused_cols == 1
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        dc = polymult.DeviceBuffer(a.nbytes)
        got = {}
        for on in (1, 0, 1):
            plan.set_fused_cols(on)
            dc.upload(np.zeros(N * L, dtype=np.uint32))
            plan.multiply_device(da, db, dc)
            polymult.synchronize()
            got[on] = dc.download((N, L))
            assert oracle.digest(got[on]) == case["c_sha"], (name, on)
        plan.set_fused_cols(1)
        dc.upload(np.zeros(N * L, dtype=np.uint32))
        plan.multiply_device_noalloc(da, db, dc)
        polymult.synchronize()
        assert oracle.digest(dc.download((N, L))) == case["c_sha"], (name, "noalloc")
        for x in (da, db, dc):
            x.close()


def test_fused_cols_batched_launches_match_gmp(golden):
    """Products of uneven K and L at N = 65536 in batched launches (the decompose
    launched once per K-step count), fused stages on, then off."""
    names = [n for n in CASES + CASES_32K if n.startswith("config-uneven-")] + ["config-n65536-w2048-01",
                                                                              "config-n32768-w868-s2"]
    cases = [golden[n] for n in names]
    plans, bufs = [], []
    for case in cases:
        q, a, b = pointgen.operands(case)
        plans.append(polymult.Plan(case["N"], case["W"], q, form=polymult.FORM_MFMA))
        bufs.append((polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b),
                     polymult.DeviceBuffer(a.nbytes)))
    for on in (1, 0):
        for p in plans:
            p.set_fused_cols(on)
        for _, _, dc in bufs:
            dc.upload(np.zeros(dc.nbytes // 4, dtype=np.uint32))
        say("fused cols batch", len(plans), "on", on)
        polymult.batch_multiply_device(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
        polymult.synchronize()
        for case, (_, _, dc) in zip(cases, bufs):
            assert oracle.digest(dc.download((case["N"], case["L"]))) == case["c_sha"], (case["name"], on)
    for x in bufs:
        for y in x:
            y.close()
    for p in plans:
        p.close()


def test_fused_cols_graph_replays_match_gmp(golden):
    """The 16-product batch of N = 65536, 2048-bit q as a captured graph with one
    run per product and as chains, replayed twice each."""
    names = ["config-n65536-w2048-%02d" % i for i in range(16)]
    cases = [golden[n] for n in names]
    plans, bufs = [], []
    for case in cases:
        q, a, b = pointgen.operands(case)
        plans.append(polymult.Plan(case["N"], case["W"], q, form=polymult.FORM_MFMA))
        bufs.append((polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b),
                     polymult.DeviceBuffer(a.nbytes)))
    for shape in ((16, 16, -1, -1), (1, 4, -1, -1), (0, 4, -1, -1)):
        say("fused cols graph", shape)
        batch = polymult.Batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs],
                               shape=shape)
        for rep in range(2):
            for _, _, dc in bufs:
                dc.upload(np.zeros(dc.nbytes // 4, dtype=np.uint32))
            batch.launch()
            polymult.synchronize()
            for case, (_, _, dc) in zip(cases, bufs):
                assert oracle.digest(dc.download((case["N"], case["L"]))) == case["c_sha"], (shape, rep, case["name"])
        batch.close()
    for x in bufs:
        for y in x:
            y.close()
    for p in plans:
        p.close()


def test_fused_cols_needs_n32768_or_65536(golden):
    case = golden["config-uneven-04"]   # N = 16384
    q, _, _ = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.fused_cols == 0
        plan.set_fused_cols(False)
        with pytest.raises(polymult.PolymulError) as e:
            plan.set_fused_cols(True)
        assert e.value.code == polymult.PM_EINVAL


def test_fused_cols_environment(golden, monkeypatch):
    case = golden["config-n65536-w2048-00"]
    q, _, _ = pointgen.operands(case)
    monkeypatch.setenv("PM_FUSED_COLS", "0")
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.fused_cols == 0
    monkeypatch.setenv("PM_FUSED_COLS", "auto")
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.fused_cols == 1
    monkeypatch.setenv("PM_FUSED_COLS", "yes")
    with pytest.raises(polymult.PolymulError) as e:
        polymult.Plan(case["N"], case["W"], q)
    assert e.value.code == polymult.PM_EINVAL


@pytest.mark.parametrize("N", [65536, 32768])
@pytest.mark.parametrize("W", [2, 31, 33, 64, 65, 100, 127, 129, 255, 511, 513, 1023, 1025])
def test_fused_cols_small_limb_counts_match_oracle(N, W):
    """1 .. 33 limbs (one to three decompose K-steps, L not a multiple of 4, K
    from 18 up): the device path with the fused stages against the oracle."""
    q, a, b = oracle.gen_point(97000 + W + N, N, W)
    L = a.shape[1]
    say("fused cols small L", N, W, "L", L)
    ref = oracle.Plan(N, q).multiply(a, b)
    with polymult.Plan(N, W, q, form=polymult.FORM_MFMA) as plan:
        assert plan.fused_cols == 1
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        dc = polymult.DeviceBuffer(a.nbytes)
        plan.multiply_device(da, db, dc)
        polymult.synchronize()
        got = dc.download((N, L))
        for x in (da, db, dc):
            x.close()
    assert np.array_equal(got, ref), (W, L)
