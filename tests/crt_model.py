"""A CPU model of the matrix-core CRT kernel (k_crt_mfma, crt_tiles in
fherma-polymult_amd/csrc/pm_mfma.hip), and residues built to drive its rare paths.

The model follows the kernel step by step for one 64-coefficient tile:
  phase 1  the FP64 sums y_j / p_j and y_j M_j / q as each wave forms them
           (fused multiply-adds over primes 4 (w + 4 i) + r in that order, then
           the four waves' partials added in order, the last step contracted
           to fma(v, nq[1], sum) as the compiler emits it), v = round(...) and
           Q = floor(...), and the balanced digits of Q and v;
  phase 2  the int8 contraction, as exact integers: digit row s of
           sum_j y_j M_j - 0x80808080 sum_j M_j - Q q + v (-P mod q), folded
           four rows to a signed 64-bit limb partial;
  phase 3  four 17-limb segments with carry-in 0, the carries between them
           resolved from carry-outs, low-64-bit summaries and all-zero /
           all-ones flags (lo_summary, seg_carry), the comparison with q, the
           workgroup vote, the +-q correction and its own carry resolution.
Every intermediate is checked against the kernel's stated bounds (|G| < 2^24,
limb partials below 2^49, segment carries below 2^17, Q below 2^47), and each
rare path is counted, so that a test can show which inputs reach it.

`mutate` plants one mistake in the model (a branch switched off) so that the
tests can show that the constructed inputs detect its absence.

Reference: the CRT being restated is rns.cuh:205-238 and
rns_kernels.cuh:265-364 (x = sum_j y_j M_j + negvP_v, reduced mod q).
"""
import random
from fractions import Fraction

import numpy as np

M32 = (1 << 32) - 1
SEG = 17                 # limbs per segment (kSegMax)
ROWS = 16 * SEG          # digit rows 0 .. 271 (17 limb tiles of 16 rows)
FAR = -(1 << 31)         # kFar


def fma(a, b, c):
    """IEEE double fused multiply-add, exactly rounded."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def balanced(x, nd):
    """Balanced base-256 digits of an integer (|x| small enough for nd digits)."""
    out = []
    for _ in range(nd):
        d = ((x + 128) & 255) - 128
        out.append(d)
        x = (x - d) >> 8
    if x != 0:
        raise ValueError("value does not fit %d balanced digits" % nd)
    return out


def lo_summary(lo):
    s = lo - (1 << 64) if lo >= (1 << 63) else lo
    if lo < (1 << 30):
        return lo
    if -(1 << 30) <= s < 0:
        return s
    return FAR


class CrtModel:
    """The kernel's view of one plan: primes p, M_j = P/p_j mod q, q, -P mod q,
    and the FP64 constants the kernel reads (1/p_j, M_j/q, (-P mod q)/q)."""

    def __init__(self, p, M, q, L, negP, mq, nq1):
        self.p = [int(x) for x in p]
        self.K = len(self.p)
        self.L = L
        self.q = int(q)
        self.M = [int(x) for x in M]
        self.negP = int(negP)
        self.KS = ((self.K + 3) // 4 + 1 + 3) // 4
        assert self.KS <= 11, "beyond the kernel's K-steps"
        self.recip = [1.0 / x for x in self.p] + [0.0] * (16 * self.KS - self.K)
        self.mq = [float(x) for x in mq] + [0.0] * (16 * self.KS - self.K)
        self.nq1 = float(nq1)
        nd = 4 * L + 1
        self.mdig = np.array([balanced(m, nd) for m in self.M], dtype=np.int64)      # K x nd
        negq = balanced((1 << (32 * (L + 1))) - self.q, 4 * (L + 1) + 1)
        negq[-1] -= 1                                                                 # -q exactly
        self.negq = negq
        self.pn = balanced(self.negP, nd)
        assert sum(d << (8 * i) for i, d in enumerate(negq)) == -self.q
        corr = 0x80808080 * sum(self.M)
        self.corr = [(corr >> (32 * l)) & M32 for l in range(L + 2)]
        assert corr >> (32 * (L + 2)) == 0
        self.qlimb = [(self.q >> (32 * l)) & M32 for l in range(L)]

    # -- phase 1: FP64 sums, v, Q and their digits -----------------------------
    def estimate(self, ycol):
        """v and Q for one coefficient (ycol: its K residues)."""
        fr, fq = [], []
        for w in range(4):
            f = f2 = 0.0
            for i in range(self.KS):
                for r in range(4):
                    j = 4 * (w + 4 * i) + r
                    y = int(ycol[j]) if j < self.K else 0
                    f = fma(float(y), self.recip[j], f)
                    f2 = fma(float(y), self.mq[j], f2)
            fr.append(f)
            fq.append(f2)
        frac = ((fr[0] + fr[1]) + fr[2]) + fr[3]
        v = int(np.floor(frac + 0.5))
        v = 0 if v < 0 else min(v, self.K)
        est = fma(float(v), self.nq1, ((fq[0] + fq[1]) + fq[2]) + fq[3])
        f = np.floor(est)
        Q = 0 if f <= 0.0 else int(f)
        return v, Q, est

    # -- phase 2: exact contraction, folded to limb partials ---------------------
    def partials(self, Y, vs, Qs):
        """Y: K x n residues; returns the 68 signed limb partials per coefficient."""
        n = Y.shape[1]
        G = np.zeros((ROWS, n), dtype=np.int64)
        nd = self.mdig.shape[1]
        for j in range(self.K):
            for a in range(4):
                u = ((Y[j].astype(np.int64) >> (8 * a)) & 255) - 128
                G[a:a + nd] += self.mdig[j][:, None] * u[None, :]
        qd = np.array([balanced(Q, 6) for Q in Qs], dtype=np.int64).T     # 6 x n
        vd = np.array([balanced(v, 2) for v in vs], dtype=np.int64).T     # 2 x n
        negq = np.array(self.negq, dtype=np.int64)
        pn = np.array(self.pn, dtype=np.int64)
        for i in range(6):
            G[i:i + negq.size] += negq[:, None] * qd[i][None, :]
        for i in range(2):
            G[i:i + pn.size] += pn[:, None] * vd[i][None, :]
        assert np.abs(G).max() < (1 << 24), "int32 accumulator bound"
        part = [[0] * n for _ in range(4 * SEG)]
        for l in range(4 * SEG):
            for c in range(n):
                part[l][c] = sum(int(G[4 * l + i, c]) << (8 * i) for i in range(4))
                assert abs(part[l][c]) < (1 << 49)
        return part

    # -- phase 3: segmented carry chain, comparison, vote, correction ------------
    def prepare(self, Y):
        """Phases 1 and 2 of one tile of up to 64 coefficients (Y: K x n)."""
        assert Y.shape[1] <= 64
        est = [self.estimate(Y[:, c]) for c in range(Y.shape[1])]
        vs, Qs = [e[0] for e in est], [e[1] for e in est]
        assert max(Qs) < (1 << 47)
        return vs, Qs, self.partials(Y, vs, Qs)

    def tile(self, prep, mutate=(), hits=None):
        """Phase 3 of a prepared tile: the kernel's output per coefficient, as
        Python ints (L limbs each)."""
        _, _, part = prep
        n = len(part[0])
        hits = {} if hits is None else hits

        def hit(k):
            hits[k] = hits.get(k, 0) + 1

        lcorr = lambda l: self.corr[l] if l < self.L + 2 else 0  # noqa: E731
        qlimb = lambda l: self.qlimb[l] if l < self.L else 0     # noqa: E731
        out, signs, state = [], [], []
        for c in range(n):
            x, ex = [], []
            for w in range(4):
                seg, carry = [], 0
                for u in range(SEG):
                    s = carry + part[SEG * w + u][c] + lcorr(SEG * w + u)
                    seg.append(s & M32)
                    carry = s >> 32
                assert abs(carry) < (1 << 17)
                uz = all(v == 0 for v in seg[2:])
                uo = all(v == M32 for v in seg[2:])
                ex.append((carry, lo_summary(seg[0] | (seg[1] << 32)), uz, uo))
                x.append(seg)
            C, cin = 0, []
            for w in range(4):
                cin.append(C)
                cout, lo_s, uz, uo = ex[w]
                adj = 0
                if C > 0 and uo and lo_s != FAR and lo_s < 0 and lo_s + C >= 0:
                    adj = 1
                    hit("seg_carry_up")
                    if "no_seg_carry_up" in mutate:
                        adj = 0
                if C < 0 and uz and lo_s != FAR and 0 <= lo_s < -C:
                    adj = -1
                    hit("seg_carry_down")
                    if "no_seg_carry_down" in mutate:
                        adj = 0
                C = cout + adj
            ctop = C
            for w in range(4):
                C = cin[w]
                for u in range(SEG):
                    t = x[w][u] + C
                    x[w][u] = t & M32
                    C = t >> 32
            cmp = 0
            for w in (3, 2, 1, 0):
                cw = 0
                for u in range(SEG - 1, -1, -1):
                    ql = qlimb(SEG * w + u)
                    d = 1 if x[w][u] > ql else (-1 if x[w][u] < ql else 0)
                    cw = cw if cw != 0 else d
                cmp = cmp if cmp != 0 else cw
            sign = 1 if ctop < 0 else (-1 if cmp >= 0 else 0)
            if sign == 1:
                hit("q_high")          # R < 0: the quotient estimate was one too large
                if "no_add_q" in mutate:
                    sign = 0
            elif sign == -1:
                hit("q_low")           # R >= q: one too small
                if "no_sub_q" in mutate:
                    sign = 0
            signs.append(sign)
            state.append(x)
        # the workgroup's vote: the correction runs when any lane needs it
        lanes = [0] if "vote_lane0_only" in mutate else range(n)
        needed = any(signs[c] != 0 for c in lanes)
        if "no_correction" in mutate:
            needed = False
        hit("vote_correct" if needed else "vote_skip")
        for c in range(n):
            x, sign = state[c], signs[c]
            if needed:
                ex = []
                for w in range(4):
                    cc, az, ao = 0, True, True
                    for u in range(SEG):
                        t = x[w][u] + sign * qlimb(SEG * w + u) + cc
                        x[w][u] = t & M32
                        cc = t >> 32
                        az = az and x[w][u] == 0
                        ao = ao and x[w][u] == M32
                    assert abs(cc) <= 1
                    ex.append((cc, az, ao))
                for wave in range(4):
                    C = 0
                    for w in range(3):
                        cc, az, ao = ex[w]
                        over = 0
                        if C == 1 and ao:
                            over = 1
                            if wave == 3:
                                hit("carry_over_ones")
                            if "no_over_up" in mutate:
                                over = 0
                        elif C == -1 and az:
                            over = -1
                            if wave == 3:
                                hit("carry_over_zeros")
                            if "no_over_down" in mutate:
                                over = 0
                        C = cc + over if w < wave else C
                    for u in range(SEG):
                        t = x[wave][u] + C
                        x[wave][u] = t & M32
                        C = t >> 32
            limbs = [x[l // SEG][l % SEG] for l in range(self.L)]
            out.append(sum(v << (32 * i) for i, v in enumerate(limbs)))
        return out


# -- residues built to reach the rare paths --------------------------------------

def residues_of(values, p):
    """y_j = c (P/p_j)^-1 mod p_j for each signed integer c: the inverse NTT's
    output for a product whose exact coefficient is c (rns.cuh:205-238)."""
    P = 1
    for x in p:
        P *= int(x)
    inv = [pow((P // int(x)) % int(x), -1, int(x)) for x in p]
    Y = np.zeros((len(p), len(values)), dtype=np.uint32)
    for j, x in enumerate(p):
        x = int(x)
        Y[j] = [(c % x) * inv[j] % x for c in values]
    return Y


def targets(q, L, P, n, seed, margin_bits=22):
    """n signed coefficients c = m q + d, |c| < P / 2^margin_bits, with m random
    and d cycling through values that put T / q next to an integer from either
    side (d = 0, +-1, +-2: the FP64 quotient one low or one high) and values
    whose remainder has whole 17-limb segments of zeros or ones (powers of two
    at the segment boundaries, and q minus them). A few m are 0 and +-1."""
    rng = random.Random(seed)
    mmax = P // (q << margin_bits)
    seg = 32 * SEG
    ds = [0, 1, -1, 2, -2, q - 1, q - 2]
    for k in (1, 2, 3):
        if seg * k < 32 * L:
            ds += [1 << (seg * k), (1 << (seg * k)) - 1, q - (1 << (seg * k)), -(1 << (seg * k))]
    out = []
    for i in range(n):
        d = ds[i % len(ds)]
        m = (i % 3) - 1 if i % 29 == 3 else rng.randint(-mmax, mmax)
        c = m * q + d
        if abs(c) >= P >> margin_bits:
            c = d
        out.append(c)
    return out


def random_values(P, n, seed, margin_bits=22):
    """n signed coefficients uniform in (-P / 2^margin_bits, P / 2^margin_bits):
    ordinary products, whose quotient estimates are almost always exact."""
    rng = random.Random(seed)
    lim = P >> margin_bits
    return [rng.randint(-lim + 1, lim - 1) for _ in range(n)]
