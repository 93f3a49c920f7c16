"""Pins the oracle (oracle/polymult_oracle.c) before anything is compared to it.

Each check is chosen so that a plausible mistake in the oracle fails one of them:
  * the seeded generator against the numpy restatement in tests/golden (a_sha, b_sha);
  * products against GMP Kronecker digests (c_sha) at every workload shape;
  * products against the reference's own solve_cpu.cpp (oracle/_ref), when built;
  * products against an exact Python-integer schoolbook for tiny N;
  * the closed forms of the reference's adversarial patterns (fuzz.py:53-63);
  * the prime rule against textbook primality and the K values of the workloads;
  * decompose against Python's big-integer remainder;
  * the forward NTT against its definition, and the convolution theorem;
  * the CRT against a reconstruction from residues of known integers;
  * the margin: sum_j y_j / p_j stays within 2^-21 of an integer (all-(q-1) case).
"""
import random

import numpy as np
import pytest
import sympy

import oracle
import pointgen

MAX_WORK = 2 ** 27   # N * K * L bound for the CPU product checks (seconds each)


def cost(case):
    K = len(oracle.choose_primes(case["N"], int(case["q_hex"], 16).bit_length()))
    return case["N"] * K * case["L"]


def test_generator_matches_numpy_restatement(golden):
    checked = 0
    for name, case in golden.items():
        if case["N"] * case["L"] <= 2 ** 21:
            pointgen.operands(case)   # asserts q, a_sha and b_sha
            checked += 1
    assert checked > 150


@pytest.mark.parametrize("family", ["config", "pattern", "matrix"])
def test_oracle_matches_gmp_digests(golden, family):
    done = 0
    for name, case in sorted(golden.items()):
        if not name.startswith(family) or cost(case) > MAX_WORK:
            continue
        q, a, b = pointgen.operands(case)
        c = oracle.Plan(case["N"], q).multiply(a, b)
        assert oracle.digest(c) == case["c_sha"], name
        if "c_words" in case:
            assert c.reshape(-1).tolist() == case["c_words"], name
        done += 1
    assert done >= 8


def test_oracle_matches_gmp_at_largest_workload(golden):
    """One N=65536, 2048-bit product and the N=65536, 1747-bit one."""
    for name in ("config-n65536-w2048-00", "config-n65536-w1747-s1"):
        case = golden[name]
        q, a, b = pointgen.operands(case)
        c = oracle.Plan(case["N"], q).multiply(a, b)
        assert oracle.digest(c) == case["c_sha"], name


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("N,W,seed", [(2, 3, 1), (4, 64, 2), (64, 2048, 3), (1024, 868, 4), (4096, 256, 5),
                                      (8192, 1000, 6), (32768, 868, 7)])
def test_oracle_matches_reference_solve_cpu(N, W, seed):
    q, a, b = oracle.gen_point(seed, N, W)
    ref = oracle.RefSolver(N, W, q)
    try:
        assert np.array_equal(ref.run(a, b), oracle.Plan(N, q).multiply(a, b))
    finally:
        ref.close()


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_reference_solve_cpu_matches_gmp(golden):
    case = golden["config-n4096-w256-s2"]
    q, a, b = pointgen.operands(case)
    ref = oracle.RefSolver(case["N"], case["W"], q)
    try:
        assert oracle.digest(ref.run(a, b)) == case["c_sha"]
    finally:
        ref.close()


@pytest.mark.parametrize("N,W", [(2, 5), (4, 33), (8, 64), (16, 100), (32, 257), (64, 512)])
def test_oracle_matches_python_schoolbook(N, W):
    q, a, b = oracle.gen_point(1000 + N, N, W)
    qi = oracle.int_of(q)
    expect = oracle.py_negacyclic(oracle.to_ints(a), oracle.to_ints(b), qi)
    assert oracle.to_ints(oracle.Plan(N, q).multiply(a, b)) == expect
    assert oracle.to_ints(oracle.schoolbook_c(N, q, a, b)) == expect


@pytest.mark.parametrize("N,W", [(2, 31), (8, 100), (64, 868), (1024, 2048), (4096, 1747)])
def test_closed_form_patterns(N, W):
    q, a, b = oracle.gen_point(7, N, W)
    qi = oracle.int_of(q)
    plan = oracle.Plan(N, q)
    for name, (pa, pb, pc) in oracle.patterns(N, qi, oracle.to_ints(a), oracle.to_ints(b)).items():
        got = plan.multiply(oracle.to_array(pa, plan.L), oracle.to_array(pb, plan.L))
        assert oracle.to_ints(got) == pc, name
        if N <= 64:
            assert oracle.py_negacyclic(pa, pb, qi) == pc, name


def test_closed_forms_against_brute_force():
    rng = random.Random(3)
    for N in (2, 4, 8, 16, 32):
        q = rng.getrandbits(40) | 1 | (1 << 39)
        ra = [rng.randrange(q) for _ in range(N)]
        rb = [rng.randrange(q) for _ in range(N)]
        for name, (pa, pb, pc) in oracle.patterns(N, q, ra, rb).items():
            assert oracle.py_negacyclic(pa, pb, q) == pc, (N, name)


def test_prime_rule():
    """The reference's walk: every candidate p = 1 (mod step) below 2^31 is
    tested downward, none skipped, until the bits suffice."""
    pbits = 31
    for (N, W), K in zip(((4096, 256), (32768, 868), (65536, 1747), (65536, 2048)), (18, 58, 115, 134)):
        p = oracle.choose_primes(N, W)
        assert len(p) == K
        step = max(2 * N, 1 << 16)
        assert all(sympy.isprime(int(x)) and int(x) % step == 1 and x < 2 ** pbits for x in p)
        # no prime of that form lies above the first chosen one
        for c in range(int(p[0]) + step, 2 ** pbits, step):
            assert not sympy.isprime(c)
        assert all(p[i] > p[i + 1] for i in range(len(p) - 1))
        # every prime of that form between consecutive chosen ones is composite
        for hi, lo in zip(p[:-1], p[1:]):
            for c in range(int(lo) + step, int(hi), step):
                assert not sympy.isprime(c)
        need = 2 * W + N.bit_length() - 1 + 1 + 20
        bits = np.log2(p.astype(np.float64))
        assert bits.sum() > need and bits[:-1].sum() <= need


def test_plan_refusals():
    with pytest.raises(ValueError):
        oracle.Plan(3, np.array([5], dtype=np.uint32))
    with pytest.raises(ValueError):
        oracle.Plan(1 << 23, np.array([5], dtype=np.uint32))
    with pytest.raises(ValueError):
        oracle.Plan(64, np.zeros(65, dtype=np.uint32) + 1)
    with pytest.raises(ValueError):
        oracle.Plan(64, np.array([5, 0], dtype=np.uint32))   # top limb zero


def test_decompose_against_big_integers():
    N, W = 64, 2048
    q, a, _ = oracle.gen_point(9, N, W)
    plan = oracle.Plan(N, q)
    res = plan.decompose(a)
    ints = oracle.to_ints(a)
    for j, p in enumerate(plan.primes):
        assert res[j].tolist() == [x % int(p) for x in ints]


def test_forward_ntt_against_definition():
    N, W = 32, 100
    q, a, _ = oracle.gen_point(10, N, W)
    plan = oracle.Plan(N, q)
    res = plan.decompose(a)
    fwd = plan.ntt_forward(res)
    for j in range(plan.K):
        p, psi = int(plan.primes[j]), int(plan.psi[j])
        assert pow(psi, N, p) == p - 1
        assert fwd[j].tolist() == oracle.py_ntt_definition(res[j].tolist(), p, psi)


def test_convolution_theorem_and_round_trip():
    N, W = 256, 300
    q, a, b = oracle.gen_point(11, N, W)
    plan = oracle.Plan(N, q)
    A, B = plan.decompose(a), plan.decompose(b)
    y = plan.ntt_inverse(plan.pointwise(plan.ntt_forward(A), plan.ntt_forward(B)))
    ai, bi = oracle.to_ints(a), oracle.to_ints(b)
    for j in (0, plan.K // 2, plan.K - 1):
        p = int(plan.primes[j])
        conv = oracle.py_negacyclic([x % p for x in ai], [x % p for x in bi], p)
        # y_j = c (P/p_j)^-1 mod p_j
        Pj = 1
        for i, pi in enumerate(plan.primes):
            if i != j:
                Pj = Pj * int(pi) % p
        assert [v * Pj % p for v in y[j].tolist()] == conv
    # inverse after forward with the N^-1 scale folded into pointwise by a unit B
    one = np.zeros_like(b)
    one[0, 0] = 1
    rt = plan.ntt_inverse(plan.pointwise(plan.ntt_forward(A), plan.ntt_forward(plan.decompose(one))))
    for j in (0, plan.K - 1):
        p = int(plan.primes[j])
        Pj = 1
        for i, pi in enumerate(plan.primes):
            if i != j:
                Pj = Pj * int(pi) % p
        assert [v * Pj % p for v in rt[j].tolist()] == A[j].tolist()


def test_crt_reconstructs_known_integers():
    N, W = 16, 640
    q, _, _ = oracle.gen_point(12, N, W)
    plan = oracle.Plan(N, q)
    qi = oracle.int_of(q)
    P = 1
    for p in plan.primes:
        P *= int(p)
    rng = random.Random(5)
    bound = N * qi * qi
    cs = [rng.randrange(-bound, bound) for _ in range(N - 2)] + [bound - 1, -bound + 1]
    Y = np.zeros((plan.K, N), dtype=np.uint32)
    for j, p in enumerate(plan.primes):
        p = int(p)
        inv = pow(P // p % p, -1, p)
        Y[j] = [(c % p) * inv % p for c in cs]
    c, worst = plan.crt(Y)
    assert oracle.to_ints(c) == [x % qi for x in cs]
    assert worst < 2.0 ** -21


def test_margin_holds_at_worst_case(golden):
    case = golden["pattern-all_q_minus_1-n4096-w256"]
    q, a, b = pointgen.operands(case)
    plan = oracle.Plan(case["N"], q)
    c = plan.multiply(a, b)
    assert oracle.digest(c) == case["c_sha"]
    assert plan.last_margin < 2.0 ** -21


def test_checker_catches_a_flipped_bit(golden):
    """A single flipped bit in the oracle's output must change the digest."""
    case = golden["config-n4096-w256-s1"]
    q, a, b = pointgen.operands(case)
    c = oracle.Plan(case["N"], q).multiply(a, b)
    assert oracle.digest(c) == case["c_sha"]
    c[1234, 3] ^= 1 << 17
    assert oracle.digest(c) != case["c_sha"]
