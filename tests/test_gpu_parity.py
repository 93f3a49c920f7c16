"""GPU parity: the HIP library against the oracle and the gmpy2 digests.

Every product is compared bit for bit: with the committed sha256 of the exact
product (tests/golden/golden.json, computed with GMP) and, step by step, with
the oracle's C restatement of the reference's algorithm on the same inputs.
"""
import random
import sys

import numpy as np
import pytest

import oracle
import pointgen

polymult = pytest.importorskip("polymult")

pytestmark = pytest.mark.gpu

FORMS = [("plain", 0), ("mfma", 1)]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    name, cus = polymult.device_info()
    say("device", name, "CUs", cus)
    yield
    polymult.synchronize()


SMALL = ["matrix-n2-w3-random", "matrix-n2-w32-prime", "matrix-n2-w33-random", "matrix-n4-w30-prime",
         "matrix-n4-w64-random", "matrix-n8-w97-prime", "matrix-n16-w127-random", "matrix-n32-w31-prime",
         "matrix-n64-w2048-random", "matrix-n128-w1025-prime", "matrix-n256-w65-random",
         "matrix-n512-w257-prime", "matrix-n1024-w868-random", "matrix-n2048-w129-prime"]
CONFIGS = ["config-n4096-w256-s1", "config-n4096-w256-s2", "config-n32768-w868-s1", "config-n65536-w1747-s1",
           "config-n65536-w2048-00", "matrix-n8192-w868-prime", "matrix-n16384-w868-random",
           "matrix-n131072-w512-prime", "matrix-n4096-w2048-random"]
PATTERNS = ["pattern-%s-n%d-w%d" % (p, n, w) for p in ("all_q_minus_1", "zero", "wrap", "identity", "alternating")
            for n, w in ((4, 64), (64, 100), (256, 1747), (4096, 256), (32768, 868), (65536, 2048))]


def run_case(golden, name, form):
    case = golden[name]
    q, a, b = pointgen.operands(case)
    say("case", name, "N", case["N"], "L", case["L"], "form", form)
    with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
        c = plan.multiply(a, b)
    got = oracle.digest(c)
    if got != case["c_sha"]:
        ref = oracle.Plan(case["N"], q).multiply(a, b)
        bad = np.nonzero(np.any(ref != c, axis=1))[0]
        pytest.fail("%s: %d of %d coefficients differ, first %s" % (name, bad.size, case["N"], bad[:5]))


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
@pytest.mark.parametrize("name", SMALL + CONFIGS)
def test_product_matches_gmp(golden, name, form):
    run_case(golden, name, form)


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
@pytest.mark.parametrize("name", PATTERNS)
def test_patterns_match_gmp(golden, name, form):
    run_case(golden, name, form)


# The reference's test-matrix points above N = 2^17 (run_tests.sh:18-21), up to the
# largest N = 2^22, with random odd q, the spec's prime q and all-(q-1) operands:
# the generic one-barrier-per-stage NTT kernels at their largest grids.
LARGEST = ["matrix-n%d-w%d-%s" % (n, w, kind) for n, w in ((262144, 128), (1048576, 31), (4194304, 32))
           for kind in ("random", "prime", "prime-allqm1")]


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
@pytest.mark.parametrize("name", LARGEST)
def test_largest_sizes_match_gmp(golden, name, form):
    run_case(golden, name, form)


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
def test_whole_test_matrix_matches_gmp(golden, form):
    """Every point of the reference's test matrix (run_tests.sh:18-21 and the
    spec's limb edges), each with a random odd q, the spec's prime q and the
    all-(q-1) operands: 141 products, each against its GMP digest."""
    names = sorted(n for n in golden if n.startswith("matrix-"))
    bad = []
    for name in names:
        case = golden[name]
        q, a, b = pointgen.operands(case)
        say("matrix", name, "form", form)
        with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
            if oracle.digest(plan.multiply(a, b)) != case["c_sha"]:
                bad.append(name)
    assert len(names) == 141 and not bad, bad


@pytest.mark.parametrize("seed", range(24))
def test_fuzz_matches_oracle(seed):
    """Random points in the spirit of the reference's fuzz runs (fuzz.py): N a
    random power of two up to 2^14, q a random odd integer of 2 .. 2048 bits,
    a and b uniform below q; the plain and the matrix-core form against the
    oracle."""
    rng = random.Random(91000 + seed)
    N, W = 1 << rng.randint(1, 14), rng.randint(2, 2048)
    q, a, b = oracle.gen_point(91000 + seed, N, W)
    say("fuzz", seed, "N", N, "W", W)
    ref = oracle.Plan(N, q).multiply(a, b)
    for form in (polymult.FORM_PLAIN, polymult.FORM_MFMA):
        with polymult.Plan(N, W, q, form=form) as plan:
            assert np.array_equal(plan.multiply(a, b), ref), (seed, N, W, form)


def _limbs(x, L):
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


@pytest.mark.parametrize("seed", range(24))
def test_fuzz_extreme_operands_match_oracle(seed):
    """Random N and q width with the operands at the top of their range: q of
    the form 2^W - 1 or 2^(W-1) + 1 with a = q - 1 and b random (even seeds),
    or a random odd q with a = b = q - 1 (odd seeds), so that every CRT sum
    sits near its largest value; both forms against the oracle."""
    rng = random.Random(92000 + seed)
    N, W = 1 << rng.randint(1, 14), rng.randint(2, 2048)
    q, _, _ = oracle.gen_point(92000 + seed, N, W)
    L = q.shape[0]
    if seed % 2 == 0:
        qi = (1 << W) - 1 if rng.random() < 0.5 else (1 << (W - 1)) + 1
        q = _limbs(qi, L)
        a = np.tile(_limbs(qi - 1, L), (N, 1))
        b = np.stack([_limbs(rng.randrange(qi), L) for _ in range(N)])
    else:
        qi = sum(int(w) << (32 * i) for i, w in enumerate(q))
        a = np.tile(_limbs(qi - 1, L), (N, 1))
        b = a.copy()
    say("fuzz-extreme", seed, "N", N, "W", W)
    ref = oracle.Plan(N, q).multiply(a, b)
    for form in (polymult.FORM_PLAIN, polymult.FORM_MFMA):
        with polymult.Plan(N, W, q, form=form) as plan:
            assert np.array_equal(plan.multiply(a, b), ref), (seed, N, W, form)


@pytest.mark.parametrize("N,W", [(64, 100), (256, 1747), (4096, 256), (65536, 2048), (131072, 512)])
def test_plan_primes(N, W):
    """A plan takes the reference's primes (the reference's K)."""
    q, _, _ = oracle.gen_point(11, N, W)
    with polymult.Plan(N, W, q) as plan:
        assert np.array_equal(plan.primes, oracle.choose_primes(N, _qbits(q)))


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
@pytest.mark.parametrize("N,W,seed,ntt", [(64, 100, 3, 0), (4096, 256, 5, 0), (8192, 868, 6, 0), (32768, 868, 7, 0),
                                          (65536, 2048, 8, 0), (65536, 2048, 8, 1), (65536, 1747, 9, 1),
                                          (65536, 64, 10, 1)])
def test_stages_match_oracle(N, W, seed, ntt, form):
    """Each of the five steps, device to device, against the oracle's step (the
    stage entry points leave canonical residues); at N = 65536 with the column
    passes in both forms (butterflies, int8 matrix cores)."""
    say("stages", N, W, seed, "form", form, "ntt", ntt)
    q, a, b = oracle.gen_point(seed, N, W)
    ref = oracle.Plan(N, q)
    with polymult.Plan(N, W, q, form=form) as plan:
        plan.set_ntt_form(ntt)
        assert plan.ntt_form == ntt
        assert plan.K == ref.K and np.array_equal(plan.primes, ref.primes)
        K, L = plan.K, plan.L
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        A = polymult.DeviceBuffer(4 * K * N)
        B = polymult.DeviceBuffer(4 * K * N)
        C = polymult.DeviceBuffer(4 * N * L)

        plan.stage_decompose(da, A)
        plan.stage_decompose(db, B)
        rA, rB = ref.decompose(a), ref.decompose(b)
        assert np.array_equal(A.download((K, N)), rA), "decompose(a)"
        assert np.array_equal(B.download((K, N)), rB), "decompose(b)"

        plan.stage_ntt(A, True)
        plan.stage_ntt(B, True)
        fA, fB = ref.ntt_forward(rA), ref.ntt_forward(rB)
        assert np.array_equal(A.download((K, N)), fA), "forward NTT"
        assert np.array_equal(B.download((K, N)), fB), "forward NTT"

        plan.stage_pointwise(A, B)
        pw = ref.pointwise(fA, fB)
        assert np.array_equal(A.download((K, N)), pw), "pointwise"

        plan.stage_ntt(A, False)
        y = ref.ntt_inverse(pw)
        assert np.array_equal(A.download((K, N)), y), "inverse NTT"

        plan.stage_crt(A, C)
        c_ref, _ = ref.crt(y)
        assert np.array_equal(C.download((N, L)), c_ref), "crt"

        plan.multiply_device(da, db, C)
        polymult.synchronize()
        assert np.array_equal(C.download((N, L)), c_ref), "fused device product"


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
def test_uneven_batch_matches_gmp(golden, form):
    names = sorted(n for n in golden if n.startswith("config-uneven-"))[:16]
    plans, bufs, cases = [], [], []
    for name in names:
        case = golden[name]
        q, a, b = pointgen.operands(case)
        plans.append(polymult.Plan(case["N"], case["W"], q, form=form))
        bufs.append((polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b),
                     polymult.DeviceBuffer(a.nbytes)))
        cases.append(case)
    say("batch", len(plans), "form", form)
    polymult.batch_multiply_device(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
    polymult.synchronize()
    for case, (da, db, dc), plan in zip(cases, bufs, plans):
        c = dc.download((case["N"], case["L"]))
        assert oracle.digest(c) == case["c_sha"], case["name"]
        plan.close()


@pytest.mark.parametrize("form", [f for _, f in FORMS], ids=[n for n, _ in FORMS])
def test_graph_batch_replays_match_gmp(golden, form):
    """The batch captured as a hipGraph, replayed several times with eager
    products and synchronizations in between; every replay is checked."""
    names = sorted(n for n in golden if n.startswith("config-uneven-"))[16:28]
    plans, bufs, cases = [], [], []
    for name in names:
        case = golden[name]
        q, a, b = pointgen.operands(case)
        plans.append(polymult.Plan(case["N"], case["W"], q, form=form))
        bufs.append((polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b),
                     polymult.DeviceBuffer(a.nbytes)))
        cases.append(case)
    say("graph batch", len(plans), "form", form)
    batch = polymult.Batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs])
    eager = polymult.Plan(cases[0]["N"], cases[0]["W"], pointgen.operands(cases[0])[0], form=form)
    for rep in range(3):
        for _, (_, _, dc) in zip(cases, bufs):
            dc.upload(np.zeros(dc.nbytes // 4, dtype=np.uint32))
        batch.launch()
        polymult.synchronize()
        for case, (_, _, dc) in zip(cases, bufs):
            assert oracle.digest(dc.download((case["N"], case["L"]))) == case["c_sha"], (rep, case["name"])
        # eager work between replays
        q0, a0, b0 = pointgen.operands(cases[0])
        assert oracle.digest(eager.multiply(a0, b0)) == cases[0]["c_sha"]
    batch.close()
    for p in plans:
        p.close()


def _qbits(q):
    return int(oracle.int_of(q)).bit_length()


@pytest.mark.parametrize("N,W", [(4096, 256), (32768, 868), (65536, 2048)])
def test_phase_batched_graph_matches_gmp(golden, N, W):
    """Products that share N, L and the primes captured as batched launches, one
    run per product; two replays, eager products in between, every answer vs GMP."""
    cases = sorted((c for c in golden.values() if c["N"] == N and c["W"] == W), key=lambda c: c["name"])
    plans, bufs, keep = [], [], []
    for case in cases:
        q, a, b = pointgen.operands(case)
        if _qbits(q) != W:   # other q sizes have other primes: not one batch
            continue
        plans.append(polymult.Plan(N, W, q, form=polymult.FORM_MFMA))
        bufs.append((polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b),
                     polymult.DeviceBuffer(a.nbytes)))
        keep.append(case)
    assert len(plans) >= 3
    say("phase-batched graph", len(plans), "products, N", N, "W", W)
    # one run per product on 16 branches (the timed choice may prefer chains)
    batch = polymult.Batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs],
                           shape=(len(plans), 16, -1, -1))
    assert batch.phased
    for rep in range(2):
        for _, _, dc in bufs:
            dc.upload(np.zeros(dc.nbytes // 4, dtype=np.uint32))
        batch.launch()
        polymult.synchronize()
        for case, (_, _, dc) in zip(keep, bufs):
            assert oracle.digest(dc.download((N, case["L"]))) == case["c_sha"], (rep, case["name"])
        q0, a0, b0 = pointgen.operands(keep[0])
        assert oracle.digest(plans[0].multiply(a0, b0)) == keep[0]["c_sha"]
    batch.close()
    for p in plans:
        p.close()


def test_phase_batched_chunks_match_single_products():
    """40 products (two launches of at most 32) at N = 256, W = 128, random q:
    the batched results (one run in the graph) equal the product-by-product
    ones, through the graph and through pm_batch_multiply_device on an
    explicit stream."""
    N, W, n = 256, 128, 40
    pts = [oracle.gen_point(7000 + i, N, W) for i in range(n)]
    pts = [pt for pt in pts if _qbits(pt[0]) == W]
    plans = [polymult.Plan(N, W, q, form=polymult.FORM_MFMA) for q, _, _ in pts]
    L = plans[0].L
    da = [polymult.DeviceBuffer.from_array(a) for _, a, _ in pts]
    db = [polymult.DeviceBuffer.from_array(b) for _, _, b in pts]
    ref = [polymult.DeviceBuffer(4 * N * L) for _ in pts]
    out = [polymult.DeviceBuffer(4 * N * L) for _ in pts]
    for p, a, b, c in zip(plans, da, db, ref):
        p.multiply_device(a, b, c)
    polymult.synchronize()
    want = [c.download((N, L)) for c in ref]
    say("phase-batched chunks", len(plans))
    batch = polymult.Batch(plans, da, db, out, shape=(1, 1, -1, -1))
    assert batch.phased
    batch.launch()
    polymult.synchronize()
    for i, c in enumerate(out):
        assert np.array_equal(c.download((N, L)), want[i]), i
    stream = polymult.Stream()
    for c in out:
        c.upload(np.zeros(N * L, dtype=np.uint32))
    polymult.batch_multiply_device(plans, da, db, out, stream=stream)
    stream.synchronize()
    for i, c in enumerate(out):
        assert np.array_equal(c.download((N, L)), want[i]), ("stream", i)
    batch.close()
    for p in plans:
        p.close()


@pytest.mark.parametrize("name", ["config-n4096-w256-s1", "config-n32768-w868-s1", "config-n65536-w2048-01",
                                  "matrix-n131072-w512-prime"])
def test_row_twiddle_sources_match_gmp(golden, name):
    """Both row-pass twiddle sources (full tables, generated) and both forms
    give the GMP product (the reference's fuzz runs tw-gen and tw-table)."""
    case = golden[name]
    q, a, b = pointgen.operands(case)
    for form in (polymult.FORM_PLAIN, polymult.FORM_MFMA):
        with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
            for source in (polymult.TWIDDLES_TABLE, polymult.TWIDDLES_GEN):
                plan.set_twiddles(source)
                assert plan.twiddles == source
                say("twiddles", name, "form", form, "source", source)
                assert oracle.digest(plan.multiply(a, b)) == case["c_sha"], (name, form, source)


def test_short_device_buffers_are_refused_on_the_host(golden):
    """A device buffer smaller than the plan's operand, product or residue
    array is refused with ValueError before any launch (the kernels trust the
    plan's sizes); the right sizes then give the GMP product."""
    case = golden["config-n4096-w256-s1"]
    q, a, b = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q, form=polymult.FORM_MFMA) as plan:
        n, r = plan.limb_bytes, plan.residue_bytes
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        short, C, R = polymult.DeviceBuffer(n - 4), polymult.DeviceBuffer(n), polymult.DeviceBuffer(r)
        for call in (lambda: plan.multiply_device(da, db, short), lambda: plan.multiply_device(short, db, C),
                     lambda: plan.stage_decompose(da, short), lambda: plan.stage_ntt(short, True),
                     lambda: plan.stage_crt(R, short), lambda: polymult.batch_multiply_device([plan], [da], [db], [short]),
                     lambda: polymult.batch_multiply_device([plan], [da, da], [db], [C]),
                     lambda: short.upload(a)):
            with pytest.raises(ValueError):
                call()
        plan.multiply_device(da, db, C)
        polymult.synchronize()
        assert oracle.digest(C.download((plan.N, plan.L))) == case["c_sha"]
        for buf in (da, db, short, C, R):
            buf.close()


@pytest.mark.parametrize("ntt", [0, 1], ids=["valu-ntt", "mfma-ntt"])
@pytest.mark.parametrize("name", ["config-n65536-w1747-s1", "config-n65536-w2048-00", "config-n65536-w2048-07",
                                  "pattern-all_q_minus_1-n65536-w2048", "pattern-wrap-n65536-w2048",
                                  "pattern-alternating-n65536-w2048", "matrix-n65536-w1747-prime",
                                  "matrix-n65536-w1747-prime-allqm1"])
def test_column_pass_forms_match_gmp(golden, name, ntt):
    """Products at N = 65536 with the column passes on the VALU and on the
    matrix cores, both decompose/CRT forms, against the GMP digests."""
    case = golden[name]
    q, a, b = pointgen.operands(case)
    for form in (polymult.FORM_MFMA, polymult.FORM_PLAIN):
        say("column forms", name, "ntt", ntt, "form", form)
        with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
            plan.set_ntt_form(ntt)
            assert oracle.digest(plan.multiply(a, b)) == case["c_sha"], (name, ntt, form)


def test_matrix_core_column_pass_needs_r256(golden):
    case = golden["config-n32768-w868-s1"]
    q, _, _ = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.ntt_form == polymult.NTT_VALU
        with pytest.raises(polymult.PolymulError) as e:
            plan.set_ntt_form(polymult.NTT_MFMA)
        assert e.value.code == polymult.PM_EINVAL


def test_ntt_form_default_and_environment(golden, monkeypatch):
    """New plans use the VALU column pass unless PM_NTT_FORM=mfma; an unknown
    value is refused at plan creation."""
    case = golden["config-n65536-w1747-s1"]
    q, _, _ = pointgen.operands(case)
    for env, want in ((None, polymult.NTT_VALU), ("auto", polymult.NTT_VALU), ("valu", polymult.NTT_VALU),
                      ("mfma", polymult.NTT_MFMA)):
        if env is None:
            monkeypatch.delenv("PM_NTT_FORM", raising=False)
        else:
            monkeypatch.setenv("PM_NTT_FORM", env)
        with polymult.Plan(case["N"], case["W"], q) as plan:
            assert plan.ntt_form == want, env
    monkeypatch.setenv("PM_NTT_FORM", "tensor")
    with pytest.raises(polymult.PolymulError) as e:
        polymult.Plan(case["N"], case["W"], q)
    assert e.value.code == polymult.PM_EINVAL


def test_device_sum_and_copy_yardsticks():
    """The traffic yardsticks compute what they claim: pm_device_sum is the sum
    mod 2^64 of the 32-bit words (sizes from one 16-byte unit to past one
    workgroup's grid-stride share), pm_device_copy a plain copy."""
    rng = np.random.default_rng(11)
    for words in (4, 1024, 4 * 256 * 8192 + 4, 3 << 20):
        x = rng.integers(0, 1 << 32, words, dtype=np.uint64).astype(np.uint32)
        d = polymult.DeviceBuffer.from_array(x)
        assert polymult.device_sum(d) == int(x.astype(np.uint64).sum()) % (1 << 64), words
        c = polymult.DeviceBuffer(x.nbytes)
        polymult.device_copy(c, d)
        polymult.synchronize()
        assert np.array_equal(c.download((words,)), x), words
        d.close()
        c.close()
