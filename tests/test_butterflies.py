"""The NTT kernels' butterfly word arithmetic (fherma-polymult_amd/csrc/pm_ntt.hip),
run on the CPU with 32-bit wraparound and checked against exact modular
arithmetic (x + w y, x - w y mod p for Cooley-Tukey; x + y, (x - y) w for
Gentleman-Sande) and against each form's stated output range.

The forms: Shoup twiddle pairs (w, floor(w 2^32 / p)), read from tables, and
Montgomery pairs (w' = w 2^32 mod p, w' p^-1 mod 2^32), generated in the row
pass; each canonical, or with forward residues in [0, 2p) (the reference's
primes, below 2^31). Inputs are drawn over each form's whole input range, with
its edges; a wrong range bound, a dropped correction or a sign slip fails here."""
import numpy as np
import pytest

import oracle

M32 = np.uint64(0xFFFFFFFF)


def u32(x):
    return (np.asarray(x, dtype=np.uint64) & M32).astype(np.uint64)


def mulhi(a, b):
    return (u32(a) * u32(b)) >> np.uint64(32)


def mullo(a, b):
    return u32(u32(a) * u32(b))


def mulhi_signed(a, b):
    sa = u32(a).astype(np.int64)
    sb = u32(b).astype(np.int64)
    sa = np.where(sa >= 2**31, sa - 2**32, sa)
    sb = np.where(sb >= 2**31, sb - 2**32, sb)
    return u32((sa * sb) >> 32)


def csub(r, p):
    t = u32(r - p)
    return np.minimum(t, r)


# ---- Shoup pairs -----------------------------------------------------------

def mul_shoup_lazy(a, w, ws, p):
    return u32(mullo(a, w) - mullo(mulhi(a, ws), p))


def ct(x, y, w, ws, p):
    v = csub(mul_shoup_lazy(y, w, ws, p), p)
    d = u32(x - v)
    return csub(u32(x + v), p), np.minimum(d, u32(d + p))


def gs(x, y, w, ws, p):
    return csub(u32(x + y), p), csub(mul_shoup_lazy(u32(x - y + p), w, ws, p), p)


def ct_half(x, y, w, ws, p):
    m = u32(mullo(mulhi(y, ws), p) - mullo(y, w) - 1)
    M = np.maximum(m, u32(m + p))
    xr = csub(x, p)
    return u32(~M + xr), u32(xr + M + p + 1)


# ---- Montgomery pairs ------------------------------------------------------

def tw_diff(y, wm, wq, p):
    return u32(mulhi(y, wm) - mulhi(mullo(y, wq), p))


def tw_diff_signed(d, wm, wq, p):
    return u32(mulhi_signed(d, wm) - mulhi_signed(mullo(d, wq), p))


def ct_m(x, y, wm, wq, p):
    d = tw_diff(y, wm, wq, p)
    v = np.minimum(d, u32(d + p))
    e = u32(x - v)
    return csub(u32(x + v), p), np.minimum(e, u32(e + p))


def ct_half_m(x, y, wm, wq, p):
    h1, h2 = mulhi(y, wm), mulhi(mullo(y, wq), p)
    a = u32(h2 + u32(~h1))
    M = np.maximum(a, u32(a - p))
    xr = csub(x, p)
    return u32(~M + xr), u32(xr + M + p + 1)


def gs_m(x, y, wm, wq, p):
    r = tw_diff_signed(u32(x - y), wm, wq, p)
    return csub(u32(x + y), p), np.minimum(r, u32(r + p))


# name: (function, forward, Montgomery pair, x bound, y bound, x' bound, y' bound)
# bounds are multiples of p (exclusive).
FORMS = {
    "ct": (ct, True, False, 1, 1, 1, 1),
    "gs": (gs, False, False, 1, 1, 1, 1),
    "ct_half": (ct_half, True, False, 2, 2, 2, 2),
    "ct_m": (ct_m, True, True, 1, 1, 1, 1),
    "gs_m": (gs_m, False, True, 1, 1, 1, 1),
    "ct_half_m": (ct_half_m, True, True, 2, 2, 2, 2),
}


def primes_for():
    """The largest and smallest primes of the bench point and of the smallest
    config, plus the smallest prime of the form below 2^30 (a prime far below
    the top of the word)."""
    ps = list(oracle.choose_primes(65536, 2048)[:3]) + list(oracle.choose_primes(4096, 256)[-2:])
    return [int(p) for p in ps] + [1073479681]


def draw(rng, bound, n):
    """Words in [0, bound) with both ends and the multiples of p near them."""
    edges = np.array([0, 1, 2, bound - 1, bound - 2, bound // 2, bound // 2 - 1], dtype=np.uint64)
    return np.concatenate([edges, rng.integers(0, bound, size=n, dtype=np.uint64)])


@pytest.mark.parametrize("name", sorted(FORMS))
def test_butterfly_form(name):
    fn, fwd, mont, xb, yb, xob, yob = FORMS[name]
    rng = np.random.default_rng(sum(map(ord, name)))
    n = 20000
    for p in primes_for():
        assert p < 2**31
        P = np.uint64(p)
        x = draw(rng, xb * p, n)
        y = draw(rng, yb * p, n)
        rng.shuffle(y)
        ws_all = np.concatenate([[1, p - 1, 2, (p - 1) // 2], rng.integers(1, p, size=x.size - 4)])
        w = ws_all.astype(np.uint64)
        if mont:
            wm = u32((w.astype(object) << 32) % p).astype(np.uint64)
            pinv = pow(p, -1, 2**32)
            second = mullo(wm, pinv)
            first = wm
        else:
            first = w
            second = ((w.astype(object) << 32) // p).astype(np.uint64)
        xo, yo = fn(x, y, first, second, P)
        xs, ys, wo = x.astype(object), y.astype(object), w.astype(object)
        if fwd:
            ex, ey = (xs + wo * ys) % p, (xs - wo * ys) % p
        else:
            ex, ey = (xs + ys) % p, ((xs - ys) * wo) % p
        xo_o, yo_o = xo.astype(object), yo.astype(object)
        assert np.all(xo_o % p == ex), (name, p)
        assert np.all(yo_o % p == ey), (name, p)
        assert int(xo.max()) < xob * p and int(yo.max()) < yob * p, (name, p)


def test_montgomery_pair_bounds():
    """hi(y w') - hi(m p) lies in (-p, p) for every 32-bit y and canonical w'."""
    rng = np.random.default_rng(7)
    for p in primes_for():
        y = draw(rng, 2**32, 20000)
        wm = np.concatenate([[0, p - 1, 1], rng.integers(0, p, size=y.size - 3)]).astype(np.uint64)
        wq = mullo(wm, pow(p, -1, 2**32))
        d = tw_diff(y, wm, wq, np.uint64(p)).astype(np.int64)
        d = np.where(d >= 2**31, d - 2**32, d)
        assert d.min() > -p and d.max() < p
        assert np.all((d.astype(object) - y.astype(object) * wm.astype(object) * pow(2**32, -1, p)) % p == 0)
