#!/opt/conda/bin/python3.9
"""Writes tests/golden/golden.json: seeded points and the sha256 of their exact
products, computed with GMP (gmpy2) by Kronecker substitution.

    /opt/conda/bin/python3.9 tests/golden/make_golden.py [--quick]

Independent of every other piece of this repository: the input generator
(counter-based SplitMix64) is restated here with numpy, and the product comes
from one GMP multiplication of the two polynomials packed into integers, then
folded negacyclically (the method the reference documents at make_big.py:21-29).
The tests regenerate each point with the oracle's C generator, check its digest
against a_sha / b_sha (pinning the generator), and check each product against
c_sha.

Case families:
  config-*     the workload shapes: N=4096/W=256, N=32768/W=868, N=65536/W=1747,
               the uneven 64-product batch, the 16 x (N=65536, W=2048) batch
  matrix-*     the reference's test matrix (run_tests.sh:18-21) and fuzz points
               (fuzz.py:119-129), with a random odd q and with the spec's prime q
               (largest prime below 2^W that is 1 mod 2N, oracle.py:48-57)
  pattern-*    the adversarial operand pairs of fuzz.py:53-63
Small cases (N * L <= 512) also store every limb of c.
"""
import hashlib
import json
import os
import sys

import gmpy2
import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15
H = 0xD1B54A32D192ED03


def mix64_int(z):
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def mix64_np(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def key_of(seed, stream):
    return mix64_int(seed * G + stream * H)


def words(key, idx):
    """idx: uint64 array -> uint32 words (high half of mix(key + (idx + 1) G))."""
    with np.errstate(over="ignore"):
        z = np.uint64(key) + (idx + np.uint64(1)) * np.uint64(G)
    return (mix64_np(z) >> np.uint64(32)).astype(np.uint32)


def top_mask(W):
    r = W % 32
    return (1 << r) - 1 if r else 0xFFFFFFFF


def gen_q(seed, W):
    L = (W + 31) // 32
    q = words(key_of(seed, 0), np.arange(L, dtype=np.uint64))
    q[L - 1] &= np.uint32(top_mask(W))
    q[L - 1] |= np.uint32(1 << ((W - 1) % 32))
    q[0] |= np.uint32(1)
    return q


def less_than(x, q):
    """Row-wise x < q for (n, L) limbs."""
    lt = np.zeros(x.shape[0], dtype=bool)
    eq = np.ones(x.shape[0], dtype=bool)
    for l in range(x.shape[1] - 1, -1, -1):
        lt |= eq & (x[:, l] < q[l])
        eq &= x[:, l] == q[l]
    return lt


def gen_operand(seed, stream, q, W, N):
    L = (W + 31) // 32
    key = key_of(seed, stream)
    out = np.zeros((N, L), dtype=np.uint32)
    pending = np.arange(N, dtype=np.uint64)
    lanes = np.arange(L, dtype=np.uint64)
    for t in range(64):
        if pending.size == 0:
            break
        idx = ((pending * np.uint64(64) + np.uint64(t)) * np.uint64(L))[:, None] + lanes[None, :]
        cand = words(key, idx)
        cand[:, L - 1] &= np.uint32(top_mask(W))
        ok = less_than(cand, q)
        out[pending[ok].astype(np.int64)] = cand[ok]
        pending = pending[~ok]
    assert pending.size == 0
    return out


def to_int(limbs):
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u4").tobytes(), "little")


def ints_to_limbs(vals, L):
    return np.frombuffer(b"".join(int(v).to_bytes(4 * L, "little") for v in vals), dtype="<u4").reshape(len(vals), L)


def kronecker(a, b, q, N):
    """Exact negacyclic product mod q of two (N, L) limb arrays."""
    L = a.shape[1]
    slot = max((2 * q.bit_length() + N.bit_length() + 1 + 7) // 8, 4 * L)

    def pack(x):
        buf = np.zeros((N, slot), dtype=np.uint8)
        buf[:, :4 * L] = np.ascontiguousarray(x, dtype="<u4").view(np.uint8).reshape(N, 4 * L)
        return gmpy2.mpz(int.from_bytes(buf.tobytes(), "little"))

    prod = pack(a) * pack(b)
    raw = int(prod).to_bytes(2 * N * slot, "little")
    qz = gmpy2.mpz(q)
    out = []
    for k in range(N):
        lo = gmpy2.mpz(int.from_bytes(raw[k * slot:(k + 1) * slot], "little"))
        hi = gmpy2.mpz(int.from_bytes(raw[(k + N) * slot:(k + N + 1) * slot], "little"))
        out.append(int((lo - hi) % qz))
    return ints_to_limbs(out, L)


def schoolbook(a, b, q):
    N = len(a)
    c = [0] * N
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            if i + j < N:
                c[i + j] += x * y
            else:
                c[i + j - N] -= x * y
    return [v % q for v in c]


def spec_prime(N, W):
    """Largest prime below 2^W that is 1 mod 2N (oracle.py:48-57's rule)."""
    m = 2 * N
    q = (1 << W) - 1
    q -= (q - 1) % m
    while q > 1:
        if gmpy2.is_prime(q, 50):
            return q
        q -= m
    raise RuntimeError("no prime")


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def pattern_operands(name, N, q, L, ra, rb):
    qm = q - 1
    if name == "random":
        return ra, rb
    if name == "all_q_minus_1":
        v = ints_to_limbs([qm] * N, L)
        return v, v
    if name == "zero":
        return np.zeros((N, L), dtype=np.uint32), rb
    if name == "wrap":
        return ints_to_limbs([0] * (N - 1) + [1], L), ints_to_limbs([0] * (N - 1) + [qm], L)
    if name == "identity":
        return ra, ints_to_limbs([1] + [0] * (N - 1), L)
    if name == "alternating":
        return (ints_to_limbs([qm * (i % 2) for i in range(N)], L),
                ints_to_limbs([qm * ((i + 1) % 2) for i in range(N)], L))
    raise KeyError(name)


def case(name, N, W, seed, q_kind="random", pattern="random"):
    L = (W + 31) // 32
    if q_kind == "random":
        qa = gen_q(seed, W)
        q = to_int(qa)
    else:
        q = spec_prime(N, W)
        qa = ints_to_limbs([q], L)[0]
    ra = gen_operand(seed, 1, qa, W, N)
    rb = gen_operand(seed, 2, qa, W, N)
    a, b = pattern_operands(pattern, N, q, L, ra, rb)
    c = kronecker(a, b, q, N)
    if N <= 64:
        assert [to_int(r) for r in c] == schoolbook([to_int(r) for r in a], [to_int(r) for r in b], q), name
    rec = {"name": name, "N": N, "W": W, "L": L, "seed": seed, "q_kind": q_kind, "pattern": pattern,
           "q_hex": "%x" % q, "a_sha": sha(a), "b_sha": sha(b), "c_sha": sha(c)}
    if N * L <= 512:
        rec["c_words"] = [int(v) for v in c.reshape(-1)]
    print("  %-40s N=%-7d W=%-5d L=%-3d %s" % (name, N, W, L, rec["c_sha"][:16]), flush=True)
    return rec


def batch_shapes(seed, count):
    """The uneven batch: N in {2^12..2^16} and W in [128, 2048], from stream 7."""
    key = key_of(seed, 7)
    w = words(key, np.arange(2 * count, dtype=np.uint64))
    return [(1 << (12 + int(w[2 * i]) % 5), 128 + int(w[2 * i + 1]) % 1921) for i in range(count)]


def main():
    quick = "--quick" in sys.argv
    cases = []
    print("configs", flush=True)
    for s in (1, 2, 3):
        cases.append(case("config-n4096-w256-s%d" % s, 4096, 256, s))
    for s in (1, 2):
        cases.append(case("config-n32768-w868-s%d" % s, 32768, 868, s))
    cases.append(case("config-n65536-w1747-s1", 65536, 1747, 1))
    if not quick:
        for i, (N, W) in enumerate(batch_shapes(4000, 64)):
            cases.append(case("config-uneven-%02d" % i, N, W, 4000 + i))
        for i in range(16):
            cases.append(case("config-n65536-w2048-%02d" % i, 65536, 2048, 5000 + i))
    print("patterns", flush=True)
    for N, W in ((4, 64), (64, 100), (256, 1747), (4096, 256), (32768, 868), (65536, 2048)):
        for pat in ("all_q_minus_1", "zero", "wrap", "identity", "alternating"):
            cases.append(case("pattern-%s-n%d-w%d" % (pat, N, W), N, W, 77, pattern=pat))
    print("matrix", flush=True)
    small = [(4, 30), (8, 60), (16, 64), (64, 32), (64, 100), (256, 128), (512, 250), (1024, 512),
             (1024, 868), (2048, 868), (256, 1747), (128, 2048)]
    big = [(2048, 129), (4096, 868), (8192, 256), (8192, 868), (16384, 868), (4096, 1024), (32768, 868),
           (65536, 1747), (131072, 512)]
    fuzz_fixed = [(2, 3), (2, 32), (2, 33), (4, 64), (4, 65), (8, 96), (8, 97), (16, 127), (32, 31),
                  (256, 63), (256, 64), (256, 65), (512, 255), (512, 256), (512, 257), (128, 1023),
                  (128, 1024), (128, 1025), (64, 1999), (64, 2047), (64, 2048), (4096, 2048),
                  (1 << 17, 64), (1 << 18, 128)]
    if not quick:
        fuzz_fixed += [(1 << 20, 31), (1 << 22, 32)]
    for N, W in small + big + fuzz_fixed:
        cases.append(case("matrix-n%d-w%d-random" % (N, W), N, W, 11))
        cases.append(case("matrix-n%d-w%d-prime" % (N, W), N, W, 11, q_kind="spec_prime"))
        cases.append(case("matrix-n%d-w%d-prime-allqm1" % (N, W), N, W, 11, q_kind="spec_prime",
                          pattern="all_q_minus_1"))
    doc = {"generator": "splitmix64-counter-v1", "gmp": gmpy2.mp_version(), "gmpy2": gmpy2.version(),
           "cases": cases}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0, separators=(",", ":"))
    print("wrote %d cases to %s" % (len(cases), OUT))


if __name__ == "__main__":
    main()
