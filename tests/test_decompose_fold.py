"""The matrix-core decompose's epilogue (fherma-polymult_amd/csrc/pm_mfma.hip,
mont_fold and the decompose tables of prepare_mfma), run on the CPU with 32-bit
wraparound and checked against exact integer arithmetic.

1. The Montgomery step: for 0 <= R < 2^49, t = hi(R) + mulhi(lo(R) pneg, p) +
   (lo(R) != 0) satisfies t 2^32 = R (mod p) and 0 <= t < p + 2^17 < 2p, over
   R's whole range and its edges (low word 0, all ones, the top of the range).
2. The whole epilogue: table entries T_t = f 256^t 2^32 mod p as balanced int8
   digits, the int8 sums G_b over the bytes u - 128 of random and extreme limbs,
   the pairing lo = G_0 + 256 G_1, hi = G_2 + 256 G_3, the correction constant,
   then the Montgomery step, give x f mod p for the coefficient's value x (the
   reference's decompose, rns.cuh:110-128, times the side's factor f). A dropped
   factor 2^32, a wrong correction or a carry slip fails here."""
import numpy as np
import pytest

import oracle

M32 = (1 << 32) - 1


def mont_fold(R, p, pneg):
    """The kernel's integer steps, with 32-bit wraparound (numpy uint64 arrays)."""
    R = np.asarray(R, dtype=np.uint64)
    rl = R & np.uint64(M32)
    rh = R >> np.uint64(32)
    m = (rl * np.uint64(pneg)) & np.uint64(M32)
    hi = (m * np.uint64(p)) >> np.uint64(32)
    return (rh + hi + (rl != 0).astype(np.uint64)) & np.uint64(M32)


def fold_primes():
    """The largest primes of the bench point, the smallest of the smallest
    config, and two primes of the form far below 2^31 (the fold's bounds must
    hold for any p below 2^31)."""
    return [int(x) for x in oracle.choose_primes(65536, 2048)[:3]] + \
           [int(x) for x in oracle.choose_primes(4096, 256)[-2:]] + [1073479681, 1071513601]


@pytest.mark.parametrize("p", fold_primes())
def test_montgomery_step(p):
    rng = np.random.default_rng(p)
    pneg = (-pow(p, -1, 1 << 32)) % (1 << 32)
    R = [0, 1, M32, 1 << 32, (1 << 32) + 1, (1 << 49) - 1, (1 << 49) - (1 << 32), p, 2 * p - 1, (1 << 48) + p]
    R += [int(x) << 32 for x in rng.integers(0, 1 << 17, 200)]                 # low word 0
    R += [(int(x) << 32) | M32 for x in rng.integers(0, 1 << 17, 200)]        # low word all ones
    R += [int(x) for x in rng.integers(0, 1 << 49, 20000, dtype=np.uint64)]
    t = mont_fold(np.array(R, dtype=np.uint64), p, pneg)
    for r, v in zip(R, t.tolist()):
        assert (v * (1 << 32) - r) % p == 0, (p, r, v)
        assert 0 <= v < p + (1 << 17) < 2 * p, (p, r, v)


def balanced4(x):
    out = []
    for _ in range(4):
        r = (x + 128) % 256 - 128
        out.append(r)
        x = (x - r) // 256
    assert x == 0
    return out


@pytest.mark.parametrize("p", fold_primes())
@pytest.mark.parametrize("L", [1, 8, 64])
@pytest.mark.parametrize("side", ["a", "b"])
def test_decompose_epilogue(p, L, side):
    rng = np.random.default_rng(1000 * L + p % 997)
    f = 1 if side == "a" else int(rng.integers(1, p))
    pneg = (-pow(p, -1, 1 << 32)) % (1 << 32)
    # table digits e[t][b] of T_t = f 256^t 2^32 mod p, centred, and the correction
    T = [f * pow(256, t, p) * (1 << 32) % p for t in range(4 * L)]
    e = np.array([balanced4(x if x <= p // 2 else x - p) for x in T], dtype=np.int64)   # (4L, 4)
    lift = -(-(1 << 47) // p) * p
    corr = 128 * sum(T) % p + lift
    # coefficients: random limbs, all zero, all ones, alternating extreme bytes
    n = 64
    limbs = rng.integers(0, 1 << 32, (n, L), dtype=np.uint64)
    limbs[0] = 0
    limbs[1] = M32
    limbs[2] = 0x80808080
    limbs[3] = 0x7F7F7F7F
    limbs[4, ::2] = M32
    limbs[4, 1::2] = 0
    u = np.stack([(limbs >> np.uint64(8 * k)) & np.uint64(255) for k in range(4)], axis=2).reshape(n, 4 * L)
    G = (u.astype(np.int64) - 128) @ e                                                   # (n, 4), exact
    assert np.abs(G).max() <= 1 << 22
    lo = G[:, 0] + 256 * G[:, 1]
    hi = G[:, 2] + 256 * G[:, 3]
    assert np.abs(lo).max() < 1 << 31 and np.abs(hi).max() < 1 << 31
    R = [int(a) + int(b) * 65536 + corr for a, b in zip(lo, hi)]
    assert all(0 <= r < 1 << 49 for r in R)
    t = mont_fold(np.array(R, dtype=np.uint64), p, pneg).tolist()
    for k in range(n):
        x = sum(int(w) << (32 * l) for l, w in enumerate(limbs[k]))
        assert t[k] % p == x * f % p, (p, L, side, k)
        assert t[k] < 2 * p
        assert min(t[k], (t[k] - p) % (1 << 32)) == x * f % p   # the canonical output
