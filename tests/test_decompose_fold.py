"""The matrix-core decompose's epilogue (fherma-polymult_amd/csrc/pm_mfma.hip,
decompose_tile and the decompose tables of prepare_mfma), run on the CPU with
32- and 64-bit wraparound and checked against exact integer arithmetic.

1. The Montgomery step: for 0 < R < 2^48, t = (R + m p) / 2^32 with
   m = R (-p^-1) mod 2^32 satisfies t 2^32 = R (mod p) and 0 <= t < p + 2^16 < 2p,
   over R's whole range and its edges (low word 0, all ones, the top).
2. The whole epilogue: table entries T_t = f 256^t 2^32 mod p as balanced int8
   digits, the int8 sums G_b over the bytes u - 128 of random and extreme limbs,
   the accumulators' initial values (digit 0: d_lo, digit 2: d_hi + c, with
   d = 128 sum T - 2^16 c mod p centred), the pairing lo = G_0 + 256 G_1 (int32)
   and hi = G_2 + 256 G_3 (uint32, wrapping), R = 2^16 hi + lo, then the
   Montgomery step, give x f mod p for the coefficient's value x (the
   reference's decompose, rns.cuh:110-128, times the side's factor f); and the
   int32 and uint32 intermediates stay in range for the worst inputs the table
   allows. A dropped factor 2^32, a wrong offset or a carry slip fails here."""
import numpy as np
import pytest

import oracle

M32 = (1 << 32) - 1


def mont_fold(R, p, pneg):
    """The kernel's integer steps: m = lo32(R) pneg (32-bit wrap), then
    (m p + R) >> 32 in 64 bits (numpy uint64 arrays)."""
    R = np.asarray(R, dtype=np.uint64)
    m = ((R & np.uint64(M32)) * np.uint64(pneg)) & np.uint64(M32)
    return (m * np.uint64(p) + R) >> np.uint64(32)


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
    R = [1, M32, 1 << 32, (1 << 32) + 1, (1 << 48) - 1, (1 << 48) - (1 << 32), p, 2 * p - 1, (1 << 47) + p]
    R += [int(x) << 32 for x in rng.integers(1, 1 << 16, 200)]                 # low word 0
    R += [(int(x) << 32) | M32 for x in rng.integers(0, 1 << 16, 200)]        # low word all ones
    R += [int(x) for x in rng.integers(1, 1 << 48, 20000, dtype=np.uint64)]
    t = mont_fold(np.array(R, dtype=np.uint64), p, pneg)
    for r, v in zip(R, t.tolist()):
        assert (v * (1 << 32) - r) % p == 0, (p, r, v)
        assert 0 <= v < p + (1 << 16) < 2 * p, (p, r, v)


C_OFFSET = (1 << 30) + (1 << 24)


def accumulator_inits(T, p):
    """Digit 0 and digit 2 start values, as prepare_mfma computes them."""
    d = (128 * sum(T) - (C_OFFSET << 16)) % p
    if d > p // 2:
        d -= p
    d_hi = d // 65536          # floor
    d_lo = d - 65536 * d_hi
    assert 0 <= d_lo < 65536 and abs(d_hi) <= 1 << 14
    return d_lo, d_hi + C_OFFSET


def i32(x):
    x = int(x) & M32
    return x - (1 << 32) if x >> 31 else x


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
    # table digits e[t][b] of T_t = f 256^t 2^32 mod p, centred, and the accumulator inits
    T = [f * pow(256, t, p) * (1 << 32) % p for t in range(4 * L)]
    e = np.array([balanced4(x if x <= p // 2 else x - p) for x in T], dtype=np.int64)   # (4L, 4)
    init0, init2 = accumulator_inits(T, p)
    # the worst sums the table allows (every byte at the end of its range that
    # maximizes or minimizes each digit's sum) keep the int32 / uint32 steps exact
    gmax = [int(128 * np.abs(e[:, b]).sum() + np.abs(e[:, b]).sum()) for b in range(4)]
    assert max(gmax) <= 1 << 22
    assert abs(init0) + gmax[0] + 256 * gmax[1] < 1 << 31                      # lo
    assert init2 + gmax[2] < 1 << 31                                           # digit-2 accumulator
    assert 0 < init2 - gmax[2] - 256 * gmax[3] and init2 + gmax[2] + 256 * gmax[3] < 1 << 32   # hi
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
    R = []
    for k in range(n):
        g0, g1, g2, g3 = (int(x) for x in G[k])
        lo = i32(g0 + init0 + 256 * g1)                       # int32 arithmetic, wrapping
        hi = (i32(g2 + init2) + 256 * g3) & M32               # uint32, wrapping
        assert lo == g0 + init0 + 256 * g1                     # no int32 wrap happened
        r = (hi * 65536 + lo) & ((1 << 64) - 1)               # v_mad_u64_u32 with a sign-extended addend
        assert 0 < r < 1 << 48
        assert (r - (sum(256 ** b * G[k, b] for b in range(4)) + 128 * sum(T))) % p == 0
        R.append(r)
    t = mont_fold(np.array(R, dtype=np.uint64), p, pneg).tolist()
    for k in range(n):
        x = sum(int(w) << (32 * l) for l, w in enumerate(limbs[k]))
        assert t[k] % p == x * f % p, (p, L, side, k)
        assert t[k] < 2 * p
        assert min(t[k], (t[k] - p) % (1 << 32)) == x * f % p   # the canonical output


def test_division_magic():
    """The decompose kernel's LDS fill divides an index i < 2^12 by D (L / 4 or L,
    2 <= D <= 64) as mulhi(i, floor(2^32 / D) + 1); exact for every such i and D
    (and a magic one too small would fail)."""
    i = np.arange(1 << 12, dtype=np.uint64)
    for D in range(2, 65):
        magic = np.uint64((1 << 32) // D + 1)
        assert np.array_equal((i * magic) >> np.uint64(32), i // np.uint64(D)), D
    assert not np.array_equal((i * np.uint64((1 << 32) // 3)) >> np.uint64(32), i // np.uint64(3))
