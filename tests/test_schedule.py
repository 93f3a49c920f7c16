"""CPU emulation of the register NTT schedule (fherma-polymult_amd/csrc/pm_ntt.hip).

Runs the kernels' element maps, stage order, twiddle indices and LDS placement
thread by thread in Python, checks that every stage touches each element exactly
once and every LDS slot is written once (the reference's solve_cpu.cpp does the
same for its schedule, solve_cpu.cpp:48-59), and compares the result with the
oracle's forward and inverse NTT."""
import numpy as np
import pytest

import oracle


def geometry(logN):
    logS = logN if logN <= 7 else (logN + 1) // 2
    return logS, logN - logS


def padded(u):
    return u + (u >> 4)


def ct(x, y, w, p):
    v = y * w % p
    return (x + v) % p, (x - v) % p


def gs(x, y, w, p):
    return (x + y) % p, (x - y) % p * w % p


def table_tw(tab, B):
    return lambda e, b: int(tab[(B >> (e + 1)) + b])


def gen_tw(psi, p, rho, logR, logS, N, fwd):
    """Row rho's twiddles as the kernels generate them: Z[rho][e] * Om[b]."""
    r = psi if fwd else pow(psi, p - 2, p)
    R = 1 << logR
    base = 2 * (int(format(rho, "0%db" % logR)[::-1], 2) if logR else 0) + 1
    Z = [pow(r, (base << e) % (2 * N), p) for e in range(logS)]
    Om = [pow(r, (2 * R * (int(format(b, "0%db" % (logS - 1))[::-1], 2) if logS > 1 else 0)) % (2 * N), p)
          for b in range(1 << (logS - 1))]
    return lambda e, b: Z[e] * Om[b] % p


def lds_phase_a(gen, LOGM):
    """fill_row_phase_a + LdsTwA: slot t holds stage LOGM-1-m, block t+1-2^m."""
    n = (1 << (LOGM - 4)) - 1
    tab = [None] * n
    for t in range(n):
        m = (t + 1).bit_length() - 1
        tab[t] = gen(LOGM - 1 - m, t + 1 - (1 << m))
    return lambda e, b: tab[(1 << (LOGM - 1 - e)) - 1 + b]


def phase_a(regs, LOGM, fwd, tw, p):
    GA = 1 << (LOGM - 4)
    NG = 16 // GA
    for s in range(LOGM - 4):
        e = LOGM - 1 - s if fwd else 4 + s
        d = 1 << (e - 4)
        for blk in range(GA // (2 * d)):
            w = tw(e, blk)
            for i in range(NG):
                for k in range(blk * 2 * d, blk * 2 * d + d):
                    a, b = i * GA + k, i * GA + k + d
                    regs[a], regs[b] = (ct if fwd else gs)(regs[a], regs[b], w, p)


def phase_b(regs, fwd, tw, t, p):
    for s in range(4):
        e = 3 - s if fwd else s
        d = 1 << e
        for blk in range(8 // d):
            w = tw(e, (t << (3 - e)) + blk)
            for k in range(blk * 2 * d, blk * 2 * d + d):
                regs[k], regs[k + d] = (ct if fwd else gs)(regs[k], regs[k + d], w, p)


def sub_transform(vals, LOGM, fwd, tw_a, tw_b, p):
    """One column or row through phase A / LDS / phase B exactly as a workgroup
    of M/16 threads runs it; returns the new values in element order."""
    M = 1 << LOGM
    TPS = GA = M // 16
    NG = 16 // GA
    a_map = [[t + TPS * i + 16 * k for i in range(NG) for k in range(GA)] for t in range(TPS)]
    b_map = [[16 * t + k for k in range(16)] for t in range(TPS)]
    for m in (a_map, b_map):
        assert sorted(u for t in m for u in t) == list(range(M))
    first, second = (a_map, b_map) if fwd else (b_map, a_map)
    regs = [[int(vals[u]) for u in first[t]] for t in range(TPS)]
    for t in range(TPS):
        if fwd:
            phase_a(regs[t], LOGM, True, tw_a, p)
        else:
            phase_b(regs[t], False, tw_b, t, p)
    lds = [None] * (M + M // 16)
    for t in range(TPS):
        for r, u in enumerate(first[t]):
            assert lds[padded(u)] is None
            lds[padded(u)] = regs[t][r]
    regs = [[lds[padded(u)] for u in second[t]] for t in range(TPS)]
    for t in range(TPS):
        if fwd:
            phase_b(regs[t], True, tw_b, t, p)
        else:
            phase_a(regs[t], LOGM, False, tw_a, p)
    out = [0] * M
    for t in range(TPS):
        for r, u in enumerate(second[t]):
            out[u] = regs[t][r]
    return out


def emulate(x, tw, itw, N, p, fwd, psi):
    logN = N.bit_length() - 1
    logS, logR = geometry(logN)
    S, R = 1 << logS, 1 << logR
    x = list(map(int, x))
    assert 4 <= logS <= 8 and 4 <= logR <= 8
    def cols(table, direction):
        for c in range(S):
            col = [x[c + u * S] for u in range(R)]
            t = table_tw(table, R)
            col = sub_transform(col, logR, direction, t, t, p)
            for u in range(R):
                x[c + u * S] = col[u]
    def rows(table, direction):
        for rho in range(R):
            row = x[rho * S:(rho + 1) * S]
            g = gen_tw(psi, p, rho, logR, logS, N, direction)
            assert all(g(e, b) == int(table[((N + rho * S) >> (e + 1)) + b])
                       for e in range(logS) for b in range(S >> (e + 1)))
            x[rho * S:(rho + 1) * S] = sub_transform(row, logS, direction, lds_phase_a(g, logS), g, p)
    if fwd:
        cols(tw, True)
        rows(tw, True)
    else:
        rows(itw, False)
        cols(itw, False)
    return x


@pytest.mark.parametrize("N", [256, 512, 2048, 4096])
def test_register_schedule_matches_oracle_ntt(N):
    q, a, _ = oracle.gen_point(21, N, 64)
    plan = oracle.Plan(N, q)
    res = plan.decompose(a)
    fwd = plan.ntt_forward(res)
    inv = plan.ntt_inverse(fwd)
    logN = N.bit_length() - 1
    for j in (0, plan.K - 1):
        p = int(plan.primes[j])
        psi = int(plan.psi[j])
        ipsi = pow(psi, p - 2, p)
        brv = [int(format(i, "0%db" % logN)[::-1], 2) for i in range(N)]
        tw = np.array([pow(psi, e, p) for e in brv], dtype=object)
        itw = np.array([pow(ipsi, e, p) for e in brv], dtype=object)
        assert emulate(res[j], tw, itw, N, p, True, psi) == fwd[j].tolist()
        assert emulate(fwd[j], tw, itw, N, p, False, psi) == inv[j].tolist()


def outer_network(col, tw, p):
    """decompose_cols_block's outer stages on one column of R1 values: the stage
    of distance 2^e in r takes twiddle (R1 >> (e + 1)) + blk for its block blk
    (the column pass's form with R1 rows)."""
    R1 = len(col)
    for e in reversed(range(R1.bit_length() - 1)):
        d = 1 << e
        for blk in range(R1 // (2 * d)):
            w = int(tw[(R1 >> (e + 1)) + blk])
            for k in range(blk * 2 * d, blk * 2 * d + d):
                col[k], col[k + d] = ct(col[k], col[k + d], w, p)


def fused_cols_forward(x, tw, N, p):
    """The forward NTT as a product with fused column stages runs it (N = 2^16 or
    2^15, S = 256, R1 = R / 16): the decompose kernel's stages over r
    (coefficient s + (N/R1) r, outer_network), the row pass's four stages over
    the 16 rows of each block (phase B with t = 0 and B = R + rho0, thread c
    holding column c), the exchange into the rows' phase-A layout, then the row
    pass's own eight stages (read here from the table; the generated twiddles
    are checked against it above)."""
    logN = N.bit_length() - 1
    logS, logR = geometry(logN)
    S, R = 1 << logS, 1 << logR
    R1 = R // 16
    N1 = N // R1
    x = list(map(int, x))
    for s_ in range(N1):
        col = [x[s_ + N1 * r] for r in range(R1)]
        outer_network(col, tw, p)
        for r in range(R1):
            x[s_ + N1 * r] = col[r]
    ROW = S + S // 16
    for rho0 in range(0, R, 16):
        lds = [None] * (16 * ROW)
        for c in range(S):
            col = [x[(rho0 + i) * S + c] for i in range(16)]
            phase_b(col, True, table_tw(tw, R + rho0), 0, p)
            for i in range(16):
                assert lds[i * ROW + padded(c)] is None
                lds[i * ROW + padded(c)] = col[i]
        for rl in range(16):
            row = [lds[rl * ROW + padded(u)] for u in range(S)]
            t = table_tw(tw, N + (rho0 + rl) * S)
            x[(rho0 + rl) * S:(rho0 + rl + 1) * S] = sub_transform(row, logS, True, t, t, p)
    return x


def test_outer_network_is_phase_b_for_16_rows():
    p = 2013265921
    tw = [pow(31, i * 7 + 3, p) for i in range(16)]
    a = [pow(5, i, p) for i in range(16)]
    b = list(a)
    outer_network(a, tw, p)
    phase_b(b, True, table_tw(tw, 16), 0, p)
    assert a == b


@pytest.mark.parametrize("R1", [16, 8])
def test_fused_cols_exchange_maps(R1):
    """decompose_cols_block's exchange (R1 = 16: eight waves, one column per
    thread; R1 = 8: four waves, two): a write at (batch prime 4 q + g, r, s)
    and a column read use one address pl stride + 32 r + s, the writes fill the
    buffer once, the transform steps cover the 16 primes x 32 s once; the two
    primes of a wave's step sit 32 banks apart, the four of a write 16 apart
    (mod 64)."""
    stride = R1 * 32 + 16
    waves, rqs = R1 // 2, R1 // 4
    writes = set()
    for q in range(4):
        for g in range(4):
            for w in range(waves):
                sh, rq = w // rqs, w % rqs
                for t in range(4):
                    for cl in range(16):
                        writes.add((4 * q + g) * stride + (4 * rq + t) * 32 + 16 * sh + cl)
    assert len(writes) == 16 * R1 * 32 and max(writes) < 16 * stride
    cols = []
    steps = 512 // (64 * waves)
    for it in range(steps):
        for tid in range(64 * waves):
            idx = it * 64 * waves + tid
            w, th = idx >> 6, (idx >> 5) & 1
            cols.append((4 * (w >> 1) + (w & 1) + 2 * th, idx & 31))
    assert sorted(cols) == [(pl, s_) for pl in range(16) for s_ in range(32)]
    assert {(g * stride) % 64 for g in range(4)} == {0, 16, 32, 48}
    for w in range(8):
        a, b = (4 * (w >> 1) + (w & 1)) * stride, (4 * (w >> 1) + (w & 1) + 2) * stride
        assert (b - a) % 64 == 32


@pytest.mark.parametrize("N", [65536, 32768])
def test_fused_cols_schedule_matches_oracle_ntt(N):
    q, a, _ = oracle.gen_point(22, N, 128)
    plan = oracle.Plan(N, q)
    res = plan.decompose(a)
    fwd = plan.ntt_forward(res)
    logN = N.bit_length() - 1
    brv = np.array([int(format(i, "0%db" % logN)[::-1], 2) for i in range(N)])
    j = plan.K - 1
    p = int(plan.primes[j])
    psi = int(plan.psi[j])
    pw = [1] * N
    for e in range(1, N):
        pw[e] = pw[e - 1] * psi % p
    tw = [pw[e] for e in brv]
    assert fused_cols_forward(res[j], tw, N, p) == fwd[j].tolist()
