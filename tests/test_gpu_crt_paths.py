"""The CRT kernels on residues built to reach their rare paths (GPU).

The residues come from chosen exact coefficients c = m q + d (tests/crt_model.py):
remainders next to a multiple of q from either side, so that the FP64 quotient
estimate is one low or one high and the +-q correction runs, and remainders
with whole 17-limb segments of zeros or ones, so that the segment carries of
k_crt_mfma cross them; on the CPU, tests/test_crt_model.py shows that these
inputs reach every such path of the kernel's model and that each path is
needed. Here pm_stage_crt runs them on the device in both forms (matrix-core
and plain) and every coefficient must equal c mod q, bit for bit, and the
oracle's CRT of the same residues.
"""
import sys

import numpy as np
import pytest

import oracle

from crt_model import random_values, residues_of, targets

polymult = pytest.importorskip("polymult")

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def limbs_of(values, L):
    out = np.zeros((len(values), L), dtype=np.uint32)
    for i, x in enumerate(values):
        for l in range(L):
            out[i, l] = (x >> (32 * l)) & 0xFFFFFFFF
    return out


POINTS = [
    ("sparse-2048", 4096, 2048, (1 << 2047) + 1, 4096),
    ("random-2048", 4096, 2048, None, 4096),
    ("all-ones-2048", 4096, 2048, (1 << 2048) - 1, 2048),
    ("random-1747-n65536", 65536, 1747, None, 4096),
    ("random-100", 4096, 100, None, 2048),
]


@pytest.mark.parametrize("name,N,W,qi,n", POINTS, ids=[p[0] for p in POINTS])
@pytest.mark.parametrize("form", [1, 0], ids=["mfma", "plain"])
def test_crt_paths_match_exact(name, N, W, qi, n, form):
    L = (W + 31) // 32
    if qi is None:
        qi = int(oracle.int_of(oracle.gen_point(sum(map(ord, name)), N, W)[0]))
    q = limbs_of([qi], L)[0]
    say("crt paths", name, "N", N, "W", W, "form", form)
    with polymult.Plan(N, W, q, form=form) as plan:
        p = plan.primes
        P = 1
        for x in p:
            P *= int(x)
        cs = targets(qi, L, P, n - 256, seed=n) + random_values(P, 256, seed=n + 1)
        cs += [0] * (N - n)
        Y = np.ascontiguousarray(residues_of(cs, p))
        want = limbs_of([c % qi for c in cs], L)
        ref, _ = oracle.Plan(N, q).crt(Y)
        assert np.array_equal(ref, want)
        R = polymult.DeviceBuffer.from_array(Y)
        C = polymult.DeviceBuffer(4 * N * L)
        plan.stage_crt(R, C)
        polymult.synchronize()
        got = C.download((N, L))
        bad = np.nonzero(np.any(got != want, axis=1))[0]
        assert bad.size == 0, (name, form, bad[:8])
        R.close()
        C.close()
