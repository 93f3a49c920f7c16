// Host-side plan for one point (N, q): the primes, K, the roots of unity and
// every constant the kernels read.
//
// Kept identical to the reference on purpose, because they fix K, the set of
// points a drop-in accepts and the NTT-domain values the stage API exposes:
//   * the prime walk, including the order of its FP64 sum of log2 p
//     (rns_plan.h:138-155), and the reference's error messages;
//   * the choice of psi among the primitive 2N-th roots (rns_plan.h:187-191);
//   * what each constant is (rns_plan.h:19-50): 2^(32 l) mod p_j, the
//     pointwise factor N^-1 (P/p_j)^-1 2^32 mod p_j, M_j = P/p_j mod q and
//     -vP mod q.
// How they are computed is this file's own: primality by trial division over a
// sieve, P as an exact big integer divided by each p_j, remainders mod q by
// normalized long division (Knuth, TAOCP vol. 2, 4.3.1, Algorithm D) and
// inverses by the extended Euclidean algorithm.
#include "pm_plan.h"

#include "pm_arith.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace pm {
namespace {

// --- one-word modular arithmetic --------------------------------------------

uint32_t mul_mod(uint32_t a, uint32_t b, uint32_t m) {
    return static_cast<uint32_t>(static_cast<uint64_t>(a) * b % m);
}

// base^e mod m, square-and-multiply from the top bit of e down.
uint32_t exp_mod(uint32_t base, uint64_t e, uint32_t m) {
    uint32_t acc = 1 % m;
    base %= m;
    for (int bit = e ? 63 - __builtin_clzll(e) : -1; bit >= 0; --bit) {
        acc = mul_mod(acc, acc, m);
        if ((e >> bit) & 1) acc = mul_mod(acc, base, m);
    }
    return acc;
}

// a^-1 mod m for gcd(a, m) = 1: the extended Euclidean algorithm, tracking
// only the coefficient of a.
uint32_t inv_mod(uint32_t a, uint32_t m) {
    int64_t r0 = m, r1 = a % m, s0 = 0, s1 = 1;
    while (r1 != 0) {
        const int64_t quo = r0 / r1;
        const int64_t r2 = r0 - quo * r1, s2 = s0 - quo * s1;
        r0 = r1;
        r1 = r2;
        s0 = s1;
        s1 = s2;
    }
    if (r0 != 1) throw std::logic_error("inv_mod: not invertible");
    return static_cast<uint32_t>(s0 < 0 ? s0 + m : s0);
}

// p^-1 mod 2^32 for odd p by Hensel lifting: p is its own inverse mod 8, and
// each step x (2 - p x) doubles the number of correct low bits (3, 6, 12, 24, 48).
uint32_t inv_mod_2_32(uint32_t p) {
    uint32_t x = p;
    for (int i = 0; i < 4; ++i) x *= 2u - p * x;
    return x;
}

// --- primality of the candidates (odd, below 2^31) ---------------------------

// The odd primes up to 46341 > sqrt(2^31), by a sieve of Eratosthenes built once.
const std::vector<uint32_t>& odd_primes_to_root() {
    static const std::vector<uint32_t> primes = [] {
        constexpr uint32_t kTop = 46341;
        std::vector<char> composite(kTop + 1, 0);
        std::vector<uint32_t> found;
        for (uint32_t i = 3; i <= kTop; i += 2) {
            if (composite[i]) continue;
            found.push_back(i);
            for (uint64_t k = static_cast<uint64_t>(i) * i; k <= kTop; k += 2ull * i) composite[k] = 1;
        }
        return found;
    }();
    return primes;
}

// n odd, 3 <= n < 2^31: prime when no odd prime up to sqrt(n) divides it.
bool odd_is_prime(uint32_t n) {
    for (const uint32_t d : odd_primes_to_root()) {
        if (static_cast<uint64_t>(d) * d > n) return true;
        if (n % d == 0) return false;
    }
    return true;
}

// --- naturals as little-endian u32 limbs --------------------------------------

using Limbs = std::vector<uint32_t>;

void trim(Limbs& x) {
    while (!x.empty() && x.back() == 0) x.pop_back();
}

int bits_of(const Limbs& x) {
    for (size_t i = x.size(); i-- > 0;)
        if (x[i]) return static_cast<int>(32 * i) + 32 - __builtin_clz(x[i]);
    return 0;
}

// x *= m.
void scale_by(Limbs& x, uint32_t m) {
    uint64_t carry = 0;
    for (uint32_t& w : x) {
        const uint64_t t = static_cast<uint64_t>(w) * m + carry;
        w = static_cast<uint32_t>(t);
        carry = t >> 32;
    }
    if (carry) x.push_back(static_cast<uint32_t>(carry));
}

// x mod d, and x / d into *quo if given: short division from the top limb.
uint32_t div_word(const Limbs& x, uint32_t d, Limbs* quo) {
    uint64_t rem = 0;
    if (quo) quo->assign(x.size(), 0);
    for (size_t i = x.size(); i-- > 0;) {
        const uint64_t cur = (rem << 32) | x[i];
        if (quo) (*quo)[i] = static_cast<uint32_t>(cur / d);
        rem = cur % d;
    }
    if (quo) trim(*quo);
    return static_cast<uint32_t>(rem);
}

// x mod q as q.size() limbs (q's top limb nonzero): normalized long division,
// Knuth's Algorithm D, keeping the remainder only. Both operands are shifted
// left until q's top bit is set; each quotient digit is estimated from the top
// two limbs of the running remainder, refined with q's second limb, and is
// then exact or one too large, which the add-back step repairs.
Limbs rem_by(const Limbs& x, const Limbs& q) {
    const size_t n = q.size();
    if (n == 1) return Limbs{div_word(x, q[0], nullptr)};
    const int s = __builtin_clz(q.back());
    auto shifted = [s](const Limbs& a, size_t len) {
        Limbs out(len, 0);
        for (size_t i = 0; i < a.size(); ++i) {
            out[i] |= a[i] << s;
            if (s && i + 1 < len) out[i + 1] |= a[i] >> (32 - s);
        }
        return out;
    };
    const Limbs v = shifted(q, n);
    Limbs u = shifted(x, std::max(x.size(), n) + 1);
    const uint64_t vt = v[n - 1], vs = v[n - 2];
    for (size_t j = u.size() - n; j-- > 0;) {
        const uint64_t top = (static_cast<uint64_t>(u[j + n]) << 32) | u[j + n - 1];
        uint64_t qh = top / vt, rh = top % vt;
        while ((qh >> 32) != 0 || qh * vs > ((rh << 32) | u[j + n - 2])) {
            --qh;
            rh += vt;
            if (rh >> 32) break;
        }
        // u[j .. j + n] -= qh v
        uint64_t carry = 0;
        int64_t borrow = 0;
        for (size_t i = 0; i < n; ++i) {
            const uint64_t prod = qh * v[i] + carry;
            carry = prod >> 32;
            const int64_t t = static_cast<int64_t>(u[i + j]) - static_cast<int64_t>(static_cast<uint32_t>(prod)) + borrow;
            u[i + j] = static_cast<uint32_t>(t);
            borrow = t >> 32;
        }
        const int64_t t = static_cast<int64_t>(u[j + n]) - static_cast<int64_t>(carry) + borrow;
        u[j + n] = static_cast<uint32_t>(t);
        if (t < 0) {   // the digit was one too large
            uint64_t c = 0;
            for (size_t i = 0; i < n; ++i) {
                const uint64_t a = static_cast<uint64_t>(u[i + j]) + v[i] + c;
                u[i + j] = static_cast<uint32_t>(a);
                c = a >> 32;
            }
            u[j + n] += static_cast<uint32_t>(c);
        }
    }
    // the remainder sits in u[0 .. n), shifted left by s; u[n] is zero
    Limbs r(n);
    for (size_t i = 0; i < n; ++i) r[i] = s ? (u[i] >> s) | (u[i + 1] << (32 - s)) : u[i];
    return r;
}

// q - t for t <= q, both q.size() limbs.
Limbs minus(const Limbs& q, const Limbs& t) {
    Limbs d(q.size());
    int64_t borrow = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        const int64_t x = static_cast<int64_t>(q[i]) - static_cast<int64_t>(t[i]) + borrow;
        d[i] = static_cast<uint32_t>(x);
        borrow = x >> 32;
    }
    return d;
}

// x / q in FP64 for 0 <= x < q, from the top 64 bits of each, aligned on q's top
// limb (long double keeps the quotient within one ulp of the double result).
double ratio_to(const uint32_t* x, const Limbs& q) {
    const int L = static_cast<int>(q.size());
    long double xv = 0, qv = 0;
    for (int i = L - 1; i >= std::max(0, L - 3); --i) {
        xv = xv * 4294967296.0L + x[i];
        qv = qv * 4294967296.0L + q[i];
    }
    return static_cast<double>(xv / qv);
}

// psi: the reference's root among the primitive 2N-th roots of unity mod p, the
// first g^((p - 1) / 2N), g = 2, 3, ..., whose N-th power is -1.
uint32_t root_2n(uint32_t p, uint32_t N) {
    const uint64_t e = (p - 1) / (2ull * N);
    for (uint32_t g = 2;; ++g) {
        const uint32_t c = exp_mod(g, e, p);
        if (exp_mod(c, N, p) == p - 1) return c;
    }
}

uint32_t log2_exact(uint32_t N) {
    uint32_t k = 0;
    while ((1u << k) < N) ++k;
    return k;
}

}  // namespace

void ntt_geometry(uint32_t logN, uint32_t& logS, uint32_t& logR) {
    logS = logN <= 7 ? logN : (logN + 1) / 2;
    logR = logN - logS;
}

std::vector<uint32_t> choose_primes(uint32_t N, int qbits) {
    // |c| < N q^2, and the CRT's rounding needs |c| / P below 2^-(margin + 1):
    // primes are added, largest first, until sum log2 p exceeds the bound.
    const double need_bits = 2.0 * qbits + log2_exact(N) + 1 + kMarginBits;
    const uint32_t step = std::max(2 * N, 1u << 16);
    std::vector<uint32_t> primes;
    double have_bits = 0;
    uint32_t cand = 0x7fffffffu / step * step + 1;
    while (have_bits <= need_bits) {
        if (cand < step) {
            throw std::runtime_error("not enough primes p < 2^31 with p = 1 (mod " + std::to_string(step) + ") for N = " +
                                     std::to_string(N) + ", q of " + std::to_string(qbits) + " bits");
        }
        if (odd_is_prime(cand)) {
            primes.push_back(cand);
            have_bits += std::log2(static_cast<double>(cand));
        }
        cand -= step;
    }
    if (static_cast<int>(primes.size()) > kMaxK) {
        throw std::runtime_error("K = " + std::to_string(primes.size()) + " exceeds kMaxK");
    }
    return primes;
}

HostPlan make_plan(uint32_t N, uint32_t W, const uint32_t* q, int L) {
    if (N < 2 || (N & (N - 1)) || N > (1u << kMaxLogN)) {
        throw std::runtime_error("N must be a power of two in [2, 2^" + std::to_string(kMaxLogN) + "]; got " +
                                 std::to_string(N));
    }
    if (L < 1 || L > kMaxL) {
        throw std::runtime_error("L must be in [1, " + std::to_string(kMaxL) + "]; got " + std::to_string(L));
    }
    HostPlan pl;
    pl.N = N;
    pl.W = W;
    pl.logN = log2_exact(N);
    pl.L = L;
    pl.q.assign(q, q + L);
    pl.qbits = bits_of(pl.q);
    if (pl.qbits <= 32 * (L - 1)) throw std::runtime_error("q has fewer limbs than L");
    pl.p = choose_primes(N, pl.qbits);
    const int K = pl.K = static_cast<int>(pl.p.size());
    ntt_geometry(pl.logN, pl.logS, pl.logR);

    // P exactly; M_j = P / p_j exactly, then reduced mod p_j and mod q.
    Limbs P{1};
    for (const uint32_t p : pl.p) scale_by(P, p);
    pl.pneg.resize(K);
    pl.psi.resize(K);
    pl.ipsi.resize(K);
    pl.pow2.resize(static_cast<size_t>(K) * L);
    pl.pow2b.resize(static_cast<size_t>(K) * L);
    pl.scale.resize(K);
    pl.scale_s.resize(K);
    pl.p_recip.resize(K);
    pl.M.resize(static_cast<size_t>(K) * L);
    pl.mq.resize(K);
    for (int j = 0; j < K; ++j) {
        const uint32_t p = pl.p[j];
        pl.pneg[j] = 0u - inv_mod_2_32(p);
        pl.psi[j] = root_2n(p, N);
        pl.ipsi[j] = inv_mod(pl.psi[j], p);
        pl.p_recip[j] = 1.0 / p;

        Limbs Mj;
        if (div_word(P, p, &Mj) != 0) throw std::logic_error("make_plan: p_j does not divide P");
        const uint32_t r32 = static_cast<uint32_t>((1ull << 32) % p);
        // N^-1 (P/p_j)^-1 2^32 mod p_j; the 2^32 undoes the Montgomery product's 2^-32
        pl.scale[j] = mul_mod(mul_mod(inv_mod(N % p, p), inv_mod(div_word(Mj, p, nullptr), p), p), r32, p);
        pl.scale_s[j] = shoup_companion(pl.scale[j], p);
        // 2^(32 l) mod p_j, and the same times scale_j (the b side)
        uint32_t w = 1 % p;
        for (int l = 0; l < L; ++l) {
            pl.pow2[static_cast<size_t>(j) * L + l] = w;
            pl.pow2b[static_cast<size_t>(j) * L + l] = mul_mod(w, pl.scale[j], p);
            w = mul_mod(w, r32, p);
        }
        const Limbs m = rem_by(Mj, pl.q);
        std::copy(m.begin(), m.end(), pl.M.begin() + static_cast<size_t>(j) * L);
        pl.mq[j] = ratio_to(m.data(), pl.q);
    }
    // -v P mod q for v = 0 .. K, from the exact v P
    pl.negvP.assign(static_cast<size_t>(K + 1) * L, 0);
    pl.nq.assign(K + 1, 0.0);
    for (int v = 1; v <= K; ++v) {
        Limbs vP = P;
        scale_by(vP, static_cast<uint32_t>(v));
        const Limbs t = rem_by(vP, pl.q);
        if (bits_of(t) == 0) continue;
        const Limbs neg = minus(pl.q, t);
        std::copy(neg.begin(), neg.end(), pl.negvP.begin() + static_cast<size_t>(v) * L);
        pl.nq[v] = ratio_to(neg.data(), pl.q);
    }
    return pl;
}

void build_twiddles(const HostPlan& pl, bool forward, int j0, int j1, uint32_t* out, uint32_t count) {
    const uint32_t N = pl.N;
    if (count == 0 || count > N || (count & (count - 1)) != 0) throw std::invalid_argument("twiddle count");
    std::vector<uint32_t> powers(count < N ? 0 : N);
    for (int j = j0; j < j1; ++j) {
        const uint32_t p = pl.p[j];
        const uint32_t root = forward ? pl.psi[j] : pl.ipsi[j];
        if (count == N) {
            uint32_t x = 1;
            for (uint32_t e = 0; e < N; ++e) {
                powers[e] = x;
                x = mul_mod(x, root, p);
            }
        }
        uint32_t* o = out + 2 * static_cast<size_t>(j - j0) * count;
        const uint32_t pinv = 0u - pl.pneg[j];
        for (uint32_t i = 0; i < count; ++i) {
            const uint32_t e = pl.logN ? bitrev32(i, pl.logN) : 0;
            const uint32_t w = count == N ? powers[e] : exp_mod(root, e, p);
            if (forward) {
                o[2 * i] = w;
                o[2 * i + 1] = shoup_companion(w, p);
            } else {
                const uint32_t wm = static_cast<uint32_t>((static_cast<uint64_t>(w) << 32) % p);
                o[2 * i] = wm;
                o[2 * i + 1] = wm * pinv;
            }
        }
    }
}

void build_row_twiddles(const HostPlan& pl, bool forward, int j0, int j1, uint32_t* z_pairs, uint32_t* om) {
    const uint32_t R = 1u << pl.logR, S = 1u << pl.logS;
    for (int j = j0; j < j1; ++j) {
        const uint32_t p = pl.p[j];
        const uint32_t r = forward ? pl.psi[j] : pl.ipsi[j];
        uint32_t* z = z_pairs + 2 * static_cast<size_t>(j - j0) * R * pl.logS;
        for (uint32_t rho = 0; rho < R; ++rho) {
            const uint64_t base = 2ull * (pl.logR ? bitrev32(rho, pl.logR) : 0) + 1;
            for (uint32_t e = 0; e < pl.logS; ++e) {
                // Montgomery form: the kernels' product with Om is then a Montgomery twiddle
                const uint32_t w = static_cast<uint32_t>(
                    (static_cast<uint64_t>(exp_mod(r, (base << e) % (2ull * pl.N), p)) << 32) % p);
                z[2 * (rho * pl.logS + e)] = w;
                z[2 * (rho * pl.logS + e) + 1] = shoup_companion(w, p);
            }
        }
        uint32_t* o = om + static_cast<size_t>(j - j0) * (S / 2);
        for (uint32_t b = 0; b < S / 2; ++b) {
            const uint64_t e = pl.logS > 1 ? 2ull * R * bitrev32(b, pl.logS - 1) : 0;
            o[b] = exp_mod(r, e % (2ull * pl.N), p);
        }
    }
}

void balanced_digits(const uint32_t* x, int limbs, int8_t* out, int ndigits) {
    int carry = 0;
    for (int i = 0; i < ndigits; ++i) {
        int byte = i < 4 * limbs ? static_cast<int>((x[i / 4] >> (8 * (i % 4))) & 0xFF) : 0;
        byte += carry;
        carry = 0;
        if (byte >= 128) {
            byte -= 256;
            carry = 1;
        }
        out[i] = static_cast<int8_t>(byte);
    }
    if (carry) throw std::logic_error("balanced_digits: value does not fit the digit count");
}

}  // namespace pm
