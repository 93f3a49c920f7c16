"""The library's host plan (fherma-polymult_amd/csrc/pm_plan.cpp) on the CPU.

pm_plan.cpp is compiled with g++ together with a small C shim
(tests/host_plan_shim.cpp), and every constant it builds for a point is checked
against the oracle's plan (the C restatement of rns_plan.h:118-249) and against
Python's exact integers: the primes, psi and its inverse, -p^-1 mod 2^32, the
pointwise factor, the decompose tables of both sides, M_j = P/p_j mod q,
-vP mod q and the FP64 ratios the CRT's quotient estimate uses. The points
include the long-division edge cases: one- and two-limb q, q = 2^W - 1, an
even q, and a q whose limbs below the top one are zero.
"""
import re

import numpy as np
import pytest

import oracle

import hostplan
from hostplan import limbs, plan_of, value


@pytest.fixture(scope="module")
def hp(tmp_path_factory):
    return hostplan.build(str(tmp_path_factory.mktemp("host_plan")))


POINTS = [
    ("random-n2-w3", 2, 3, None),
    ("random-n64-w100", 64, 100, None),
    ("random-n4096-w256", 4096, 256, None),
    ("random-n32768-w868", 32768, 868, None),
    ("random-n65536-w2048", 65536, 2048, None),
    ("one-limb-max", 1024, 32, (1 << 32) - 1),
    ("two-limbs-2^32+1", 256, 33, (1 << 32) + 1),
    ("all-ones-2048", 4096, 2048, (1 << 2048) - 1),
    ("even-q", 512, 300, (1 << 299) + (1 << 150) + 2),
    ("sparse-top-limb", 8192, 1025, 1 << 1024),
    ("sparse-2^2047+1", 65536, 2048, (1 << 2047) + 1),
]


@pytest.mark.parametrize("name,N,W,qi", POINTS, ids=[p[0] for p in POINTS])
def test_host_plan_constants(hp, name, N, W, qi):
    L = (W + 31) // 32
    if qi is None:
        q, _, _ = oracle.gen_point(sum(map(ord, name)), N, W)
        qi = value(q)
    q = limbs(qi, L)
    got = plan_of(hp, N, W, q)
    ref = oracle.Plan(N, q)
    K = got["K"]
    assert K == ref.K and np.array_equal(got["p"], ref.primes)
    assert np.array_equal(got["psi"], ref.psi)
    assert np.array_equal(got["M"], ref.M)
    assert np.array_equal(got["negvP"], ref.negvP)
    P = 1
    for p in got["p"]:
        P *= int(p)
    for j in range(K):
        p = int(got["p"][j])
        psi, ipsi = int(got["psi"][j]), int(got["ipsi"][j])
        assert pow(psi, N, p) == p - 1 and psi * ipsi % p == 1
        assert p * int(got["pneg"][j]) % (1 << 32) == (1 << 32) - 1
        # scale = N^-1 (P/p_j)^-1 2^32 mod p_j, and the oracle's holds it without the 2^32
        s = int(got["scale"][j])
        assert s * N * (P // p) % p == (1 << 32) % p
        assert s == int(ref.scale[j]) * (1 << 32) % p
        for l in range(L):
            w = pow(2, 32 * l, p)
            assert int(got["pow2"][j, l]) == w and int(got["pow2b"][j, l]) == w * s % p
        Mj = (P // p) % qi
        assert value(got["M"][j]) == Mj
        assert abs(got["mq"][j] - Mj / qi) <= 2.0 ** -50 * max(1.0, Mj / qi)
    for v in range(K + 1):
        t = (-v * P) % qi
        assert value(got["negvP"][v]) == t
        assert abs(got["nq"][v] - t / qi) <= 2.0 ** -50


@pytest.mark.parametrize("N,W,q,msg", [
    (3000, 64, [5, 1], "N must be a power of two"),
    (1 << 23, 64, [5, 1], "N must be a power of two"),
    (64, 64, [5, 0], "q has fewer limbs than L"),
    (1 << 22, 2048, [1] + [0] * 62 + [1 << 31], "not enough primes p < 2^31"),
])
def test_host_plan_refusals(hp, N, W, q, msg):
    """The reference's refusals and messages (rns_plan.h:120-155)."""
    with pytest.raises(ValueError, match=re.escape(msg)):
        plan_of(hp, N, W, np.array(q, dtype=np.uint32))


def test_host_plan_refuses_too_many_limbs(hp):
    with pytest.raises(ValueError, match="L must be in"):
        plan_of(hp, 64, 2080, np.ones(65, dtype=np.uint32))
