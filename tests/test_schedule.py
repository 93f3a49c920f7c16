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
