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
import ctypes
import re
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fherma-polymult_amd", "csrc")
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def hp(tmp_path_factory):
    lib = str(tmp_path_factory.mktemp("host_plan") / "libhostplan.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + CSRC,
                    os.path.join(CSRC, "pm_plan.cpp"), os.path.join(ROOT, "tests", "host_plan_shim.cpp"),
                    "-o", lib], check=True)
    L = ctypes.CDLL(lib)
    L.hp_plan.restype = ctypes.c_int
    L.hp_plan.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_int] + [_u32p] * 9 + \
                         [_f64p, _f64p, ctypes.c_char_p, ctypes.c_int]
    return L


def plan_of(hp, N, W, q):
    L = q.shape[0]
    kmax = 200
    bufs = {k: np.zeros(n, dtype=np.uint32) for k, n in
            (("p", kmax), ("pneg", kmax), ("psi", kmax), ("ipsi", kmax), ("scale", kmax),
             ("pow2", kmax * L), ("pow2b", kmax * L), ("M", kmax * L), ("negvP", (kmax + 1) * L))}
    mq, nq = np.zeros(kmax), np.zeros(kmax + 1)
    err = ctypes.create_string_buffer(256)
    u = lambda a: a.ctypes.data_as(_u32p)  # noqa: E731
    K = hp.hp_plan(N, W, u(q), L, u(bufs["p"]), u(bufs["pneg"]), u(bufs["psi"]), u(bufs["ipsi"]),
                   u(bufs["scale"]), u(bufs["pow2"]), u(bufs["pow2b"]), u(bufs["M"]), u(bufs["negvP"]),
                   mq.ctypes.data_as(_f64p), nq.ctypes.data_as(_f64p), err, 256)
    if K < 0:
        raise ValueError(err.value.decode())
    out = {k: v[:K] for k, v in bufs.items() if k in ("p", "pneg", "psi", "ipsi", "scale")}
    for k in ("pow2", "pow2b", "M"):
        out[k] = bufs[k][:K * L].reshape(K, L)
    out["negvP"] = bufs["negvP"][:(K + 1) * L].reshape(K + 1, L)
    out["mq"], out["nq"], out["K"] = mq[:K], nq[:K + 1], K
    return out


def limbs(x, L):
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


def value(row):
    return sum(int(w) << (32 * i) for i, w in enumerate(row))


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
