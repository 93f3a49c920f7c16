"""The matrix-core CRT kernel's rare paths, pinned on the CPU.

tests/crt_model.py models k_crt_mfma (pm_mfma.hip) step by step. Here it runs on
residues built from chosen exact coefficients c = m q + d (tests/crt_model.py,
`targets`) for which the FP64 quotient estimate lands one low or one high and
the remainder has whole 17-limb segments of zeros or ones, at the bench point's
width (L = 64, K = 134) with a sparse q = 2^2047 + 1 and a random q, at
N = 65536 with a 1747-bit q, at a small L, and at K = 160 primes. The model must
give c mod q for every coefficient (and the oracle's CRT on the same residues);
every rare path must be reached; and switching any one of them off in the model
must produce a wrong coefficient, which shows that the inputs test each path.
The same residues go through the kernel itself in tests/test_gpu_crt_paths.py.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle

import hostplan
from crt_model import CrtModel, random_values, residues_of, targets
from hostplan import limbs, value

PATHS = ("q_low", "q_high", "seg_carry_up", "seg_carry_down", "carry_over_ones", "carry_over_zeros",
         "vote_skip", "vote_correct")
MUTATIONS = ("no_sub_q", "no_add_q", "no_correction", "vote_lane0_only", "no_seg_carry_up", "no_seg_carry_down",
             "no_over_up", "no_over_down")


def model_point(hp, N, W, qi):
    L = (W + 31) // 32
    q = limbs(qi, L)
    h = hostplan.plan_of(hp, N, W, q)
    M = [value(r) for r in h["M"]]
    return CrtModel(h["p"], M, qi, L, value(h["negvP"][1]), h["mq"], h["nq"][1]), h


def synthetic_k160():
    """K = 160 primes (the kernel's prime-constant capacity) with a 2048-bit q:
    more primes than the rule gives any point, so the FP64 sums and the digit
    contraction run at their largest."""
    p = [int(x) for x in oracle.choose_primes(4096, 2460)]
    assert len(p) == 160
    qi = int(oracle.int_of(oracle.gen_point(160, 4096, 2048)[0]))
    P = 1
    for x in p:
        P *= x
    M = [(P // x) % qi for x in p]
    negP = (-P) % qi
    mq = [float(Fraction(m, qi)) for m in M]
    return CrtModel(p, M, qi, 64, negP, mq, float(Fraction(negP, qi))), P


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    """Every point: (model, prepared tiles, truth per tile); prepared once."""
    hp = hostplan.build(str(tmp_path_factory.mktemp("crt_model")))
    out = []
    specs = [("sparse-2048", 4096, 2048, (1 << 2047) + 1, 512),
             ("random-2048", 4096, 2048, None, 256),
             ("random-1747-n65536", 65536, 1747, None, 192),
             ("random-100", 256, 100, None, 128)]
    for name, N, W, qi, n in specs:
        if qi is None:
            qi = int(oracle.int_of(oracle.gen_point(sum(map(ord, name)), N, W)[0]))
        model, h = model_point(hp, N, W, qi)
        P = 1
        for x in h["p"]:
            P *= int(x)
        # constructed tiles, then one tile of ordinary random products
        cs = targets(qi, model.L, P, n - 64, seed=len(name)) + random_values(P, 64, seed=len(name))
        Y = residues_of(cs, h["p"])
        ref = oracle.Plan(N, limbs(qi, model.L))
        assert n <= N
        c_or, _ = ref.crt(np.ascontiguousarray(np.pad(Y, ((0, 0), (0, N - n)))))
        tiles = []
        for t0 in range(0, n, 64):
            truth = [c % qi for c in cs[t0:t0 + 64]]
            assert [value(r) for r in c_or[t0:t0 + 64]] == truth, name   # the oracle agrees
            tiles.append((model.prepare(Y[:, t0:t0 + 64]), truth))
        out.append((name, model, tiles))
    model, P = synthetic_k160()
    cs = targets(model.q, 64, P, 128, seed=160)
    Y = residues_of(cs, model.p)
    tiles = [(model.prepare(Y[:, t0:t0 + 64]), [c % model.q for c in cs[t0:t0 + 64]]) for t0 in range(0, 128, 64)]
    out.append(("k160", model, tiles))
    return out


def test_model_matches_exact_crt_and_reaches_every_path(runs):
    hits = {}
    for name, model, tiles in runs:
        for prep, truth in tiles:
            assert model.tile(prep, hits=hits) == truth, name
    missing = [p for p in PATHS if not hits.get(p)]
    print("paths reached:", hits)
    assert not missing, (missing, hits)


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_each_path_is_needed(runs, mutation):
    wrong = 0
    for _, model, tiles in runs:
        for prep, truth in tiles:
            got = model.tile(prep, mutate=(mutation,))
            wrong += sum(g != t for g, t in zip(got, truth))
    assert wrong > 0, mutation


def test_quotient_estimate_error_bound(runs):
    """|Q_est - T/q| stays far below the one unit the correction allows: the
    bound derived for the kernel (about 2^-9 at K = 160) holds with room."""
    worst = 0.0
    for name, model, _ in runs:
        if name not in ("k160", "sparse-2048"):
            continue
        P = 1
        for x in model.p:
            P *= x
        cs = targets(model.q, model.L, P, 64, seed=7)
        Y = residues_of(cs, model.p)
        for c in range(64):
            v, Q, est = model.estimate(Y[:, c])
            T = sum(int(Y[j, c]) * model.M[j] for j in range(model.K)) + v * model.negP
            worst = max(worst, abs(Fraction(est) - Fraction(T, model.q)))
    assert worst < Fraction(1, 1 << 9), float(worst)
