"""Operands for a golden case, rebuilt with the oracle's seeded generator and
its closed-form patterns; each is checked against the case's a_sha / b_sha."""
import numpy as np

import oracle


def operands(case):
    """(q limbs, a, b) for a golden case; raises if a digest disagrees."""
    N, W, L = case["N"], case["W"], case["L"]
    q_int = int(case["q_hex"], 16)
    q = oracle.to_array([q_int], L)[0]
    if case["q_kind"] == "random":
        assert np.array_equal(q, oracle.gen_q(case["seed"], W)), "generator q differs from the fixture"
    ra = oracle.gen_operand(case["seed"], 1, q, W, N)
    rb = oracle.gen_operand(case["seed"], 2, q, W, N)
    pat = case["pattern"]
    qm = q_int - 1
    if pat == "random":
        a, b = ra, rb
    elif pat == "all_q_minus_1":
        a = np.tile(oracle.to_array([qm], L), (N, 1))
        b = a.copy()
    elif pat == "zero":
        a, b = np.zeros((N, L), dtype=np.uint32), rb
    elif pat == "wrap":
        a = np.zeros((N, L), dtype=np.uint32)
        b = np.zeros((N, L), dtype=np.uint32)
        a[N - 1] = oracle.to_array([1], L)[0]
        b[N - 1] = oracle.to_array([qm], L)[0]
    elif pat == "identity":
        a = ra
        b = np.zeros((N, L), dtype=np.uint32)
        b[0, 0] = 1
    elif pat == "alternating":
        row = oracle.to_array([qm], L)[0]
        a = np.zeros((N, L), dtype=np.uint32)
        b = np.zeros((N, L), dtype=np.uint32)
        a[1::2] = row
        b[0::2] = row
    else:
        raise KeyError(pat)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    assert oracle.digest(a) == case["a_sha"], "operand a differs from the fixture"
    assert oracle.digest(b) == case["b_sha"], "operand b differs from the fixture"
    return q, a, b
