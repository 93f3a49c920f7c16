// Host-side plan: primes, K, roots, decompose and CRT constants.
// Rules follow rns_plan.h:118-249 (cited per step); the code is this build's own.
#include "pm_plan.h"

#include "pm_arith.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace pm {
namespace {

uint32_t pow_mod(uint64_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1 % p;
    b %= p;
    while (e) {
        if (e & 1) r = r * b % p;
        b = b * b % p;
        e >>= 1;
    }
    return static_cast<uint32_t>(r);
}

// Deterministic Miller-Rabin with bases 2, 3, 5, 7: exact for n < 3.2e9
// (the same test as rns_plan.h:63-83).
bool is_prime(uint32_t n) {
    if (n < 2) return false;
    for (uint32_t d : {2u, 3u, 5u, 7u})
        if (n % d == 0) return n == d;
    uint32_t d = n - 1;
    int r = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++r;
    }
    for (uint32_t a : {2u, 3u, 5u, 7u}) {
        uint64_t x = pow_mod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool witness = true;
        for (int i = 1; i < r && witness; ++i) {
            x = x * x % n;
            if (x == n - 1) witness = false;
        }
        if (witness) return false;
    }
    return true;
}

int bit_length(const uint32_t* x, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (x[i]) return 32 * i + 32 - __builtin_clz(x[i]);
    return 0;
}

bool geq(const uint32_t* x, int n, const uint32_t* q, int L) {
    for (int i = n - 1; i >= 0; --i) {
        const uint32_t qi = i < L ? q[i] : 0;
        if (x[i] != qi) return x[i] > qi;
    }
    return true;
}

void sub_q(uint32_t* x, int n, const uint32_t* q, int L) {
    uint64_t borrow = 0;
    for (int i = 0; i < n; ++i) {
        const uint64_t d = static_cast<uint64_t>(x[i]) - (i < L ? q[i] : 0) - borrow;
        x[i] = static_cast<uint32_t>(d);
        borrow = (d >> 63) & 1;
    }
}

// x -= e * q * 2^(32 shift), known to stay non-negative.
void sub_mul_q(uint32_t* x, int n, const uint32_t* q, int L, uint32_t e, int shift) {
    uint64_t carry = 0, borrow = 0;
    for (int i = shift; i < n; ++i) {
        const int k = i - shift;
        const uint64_t prod = (k < L ? static_cast<uint64_t>(e) * q[k] : 0) + carry;
        carry = prod >> 32;
        const uint64_t d = static_cast<uint64_t>(x[i]) - static_cast<uint32_t>(prod) - borrow;
        x[i] = static_cast<uint32_t>(d);
        borrow = (d >> 63) & 1;
    }
}

// x mod q in place for x < 2^64 q (n limbs): quotient estimated from the top
// limbs in double, one below the estimate so it never overshoots, then exact
// corrections.
void reduce_mod_q(uint32_t* x, int n, const HostPlan& pl) {
    double xd = 0;
    for (int i = n - 1; i >= pl.q_drop; --i) xd = xd * 4294967296.0 + x[i];
    const double f = std::floor(xd / pl.q_top);
    const uint64_t e = f >= 1 ? static_cast<uint64_t>(f) - 1 : 0;
    if (e >> 32) sub_mul_q(x, n, pl.q.data(), pl.L, static_cast<uint32_t>(e >> 32), 1);
    if (e & 0xFFFFFFFFu) sub_mul_q(x, n, pl.q.data(), pl.L, static_cast<uint32_t>(e), 0);
    while (geq(x, n, pl.q.data(), pl.L)) sub_q(x, n, pl.q.data(), pl.L);
}

// x = x * m mod q for x < q.
void mul_small_mod(std::vector<uint32_t>& x, uint32_t m, const HostPlan& pl) {
    std::vector<uint32_t> t(pl.L + 1);
    uint64_t carry = 0;
    for (int l = 0; l < pl.L; ++l) {
        const uint64_t v = static_cast<uint64_t>(x[l]) * m + carry;
        t[l] = static_cast<uint32_t>(v);
        carry = v >> 32;
    }
    t[pl.L] = static_cast<uint32_t>(carry);
    reduce_mod_q(t.data(), pl.L + 1, pl);
    for (int l = 0; l < pl.L; ++l) x[l] = t[l];
}

}  // namespace

void ntt_geometry(uint32_t logN, uint32_t& logS, uint32_t& logR) {
    logS = logN <= 7 ? logN : (logN + 1) / 2;
    logR = logN - logS;
}

std::vector<uint32_t> choose_primes(uint32_t N, int qbits, int pbits) {
    if (pbits != 30 && pbits != 31) throw std::runtime_error("the prime bound must be 2^30 or 2^31");
    uint32_t logN = 0;
    while ((1u << logN) < N) ++logN;
    // |c| < N q^2 and the CRT rounding needs |c| / P < 2^-(margin + 1)
    // (rns_plan.h:138-153).
    const double need_bits = 2.0 * qbits + logN + 1 + kMarginBits;
    double have_bits = 0;
    const uint32_t step = std::max(2 * N, 1u << 16);
    std::vector<uint32_t> primes;
    for (uint32_t c = ((1u << pbits) - 1) / step * step + 1; have_bits <= need_bits; c -= step) {
        if (c < step) {
            throw std::runtime_error("not enough primes p < 2^" + std::to_string(pbits) + " with p = 1 (mod " +
                                     std::to_string(step) +
                                     ") for N = " + std::to_string(N) + ", q of " + std::to_string(qbits) + " bits");
        }
        if (!is_prime(c)) continue;
        primes.push_back(c);
        have_bits += std::log2(static_cast<double>(c));
    }
    if (static_cast<int>(primes.size()) > kMaxK) {
        throw std::runtime_error("K = " + std::to_string(primes.size()) + " exceeds kMaxK");
    }
    return primes;
}

HostPlan make_plan(uint32_t N, uint32_t W, const uint32_t* q, int L, int pbits) {
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
    while ((1u << pl.logN) < N) ++pl.logN;
    pl.L = L;
    pl.q.assign(q, q + L);
    pl.qbits = bit_length(pl.q.data(), L);
    if (pl.qbits <= 32 * (L - 1)) throw std::runtime_error("q has fewer limbs than L");
    pl.q_drop = L > 3 ? L - 3 : 0;
    for (int i = L - 1; i >= pl.q_drop; --i) pl.q_top = pl.q_top * 4294967296.0 + pl.q[i];

    pl.p = choose_primes(N, pl.qbits, pbits);
    pl.pbits = pbits;
    const int K = pl.K = static_cast<int>(pl.p.size());

    ntt_geometry(pl.logN, pl.logS, pl.logR);

    pl.pneg.resize(K);
    pl.psi.resize(K);
    pl.ipsi.resize(K);
    pl.pow2.resize(static_cast<size_t>(K) * L);
    pl.pow2_s.resize(static_cast<size_t>(K) * L);
    pl.scale.resize(K);
    pl.scale_s.resize(K);
    pl.p_recip.resize(K);
    for (int j = 0; j < K; ++j) {
        const uint32_t p = pl.p[j];
        uint32_t inv = 1;  // p^-1 mod 2^32 by Newton's iteration (rns_plan.h:178-180)
        for (int i = 0; i < 5; ++i) inv *= 2 - p * inv;
        pl.pneg[j] = 0u - inv;
        for (int l = 0; l < L; ++l) {
            const uint32_t w = pow_mod(2, 32ull * static_cast<uint64_t>(l), p);
            pl.pow2[static_cast<size_t>(j) * L + l] = w;
            pl.pow2_s[static_cast<size_t>(j) * L + l] = shoup_companion(w, p);
        }
        // psi: the first g^((p-1)/2N), g = 2, 3, ..., whose N-th power is -1
        // (rns_plan.h:187-191). Any primitive 2N-th root gives the same product.
        uint32_t psi = 0;
        for (uint32_t g = 2; !psi; ++g) {
            const uint32_t cand = pow_mod(g, (p - 1) / (2ull * N), p);
            if (pow_mod(cand, N, p) == p - 1) psi = cand;
        }
        pl.psi[j] = psi;
        pl.ipsi[j] = pow_mod(psi, p - 2, p);
        // N^-1 (P/p_j)^-1 2^32 mod p_j (rns_plan.h:211-219).
        uint64_t Pj = 1;
        for (int i = 0; i < K; ++i)
            if (i != j) Pj = Pj * (pl.p[i] % p) % p;
        const uint64_t s = static_cast<uint64_t>(pow_mod(N, p - 2, p)) * pow_mod(Pj, p - 2, p) % p *
                           ((1ull << 32) % p) % p;
        pl.scale[j] = static_cast<uint32_t>(s);
        pl.scale_s[j] = shoup_companion(pl.scale[j], p);
        pl.p_recip[j] = 1.0 / p;
    }
    pl.pow2b.resize(static_cast<size_t>(K) * L);
    for (int j = 0; j < K; ++j)
        for (int l = 0; l < L; ++l)
            pl.pow2b[static_cast<size_t>(j) * L + l] = static_cast<uint32_t>(
                static_cast<uint64_t>(pl.scale[j]) * pl.pow2[static_cast<size_t>(j) * L + l] % pl.p[j]);

    // M_j = prod_{i != j} p_i mod q and P mod q (rns_plan.h:224-236).
    pl.M.assign(static_cast<size_t>(K) * L, 0);
    std::vector<uint32_t> Pq(L, 0);
    Pq[0] = 1;
    for (int j = 0; j < K; ++j) {
        std::vector<uint32_t> m(L, 0);
        m[0] = 1;
        for (int i = 0; i < K; ++i)
            if (i != j) mul_small_mod(m, pl.p[i], pl);
        std::copy(m.begin(), m.end(), pl.M.begin() + static_cast<size_t>(j) * L);
        mul_small_mod(Pq, pl.p[j], pl);
    }
    // negvP_v = (q - v P mod q) mod q, v = 0..K (rns_plan.h:238-247).
    pl.negvP.assign(static_cast<size_t>(K + 1) * L, 0);
    for (int v = 1; v <= K; ++v) {
        std::vector<uint32_t> t = Pq;
        mul_small_mod(t, static_cast<uint32_t>(v), pl);
        if (bit_length(t.data(), L) == 0) continue;
        std::vector<uint32_t> neg = pl.q;
        sub_q(neg.data(), L, t.data(), L);
        std::copy(neg.begin(), neg.end(), pl.negvP.begin() + static_cast<size_t>(v) * L);
    }
    // x / q in double for x < q: the top three limbs of both, aligned on q's top.
    auto ratio = [&](const uint32_t* x) {
        long double xd = 0;
        for (int i = L - 1; i >= pl.q_drop; --i) xd = xd * 4294967296.0L + x[i];
        return static_cast<double>(xd / static_cast<long double>(pl.q_top));
    };
    pl.mq.resize(K);
    for (int j = 0; j < K; ++j) pl.mq[j] = ratio(pl.M.data() + static_cast<size_t>(j) * L);
    pl.nq.resize(K + 1);
    for (int v = 0; v <= K; ++v) pl.nq[v] = ratio(pl.negvP.data() + static_cast<size_t>(v) * L);
    return pl;
}

void build_twiddles(const HostPlan& pl, bool forward, int j0, int j1, uint32_t* out) {
    const uint32_t N = pl.N;
    std::vector<uint32_t> pw(N);
    for (int j = j0; j < j1; ++j) {
        const uint32_t p = pl.p[j];
        const uint64_t r = forward ? pl.psi[j] : pl.ipsi[j];
        uint64_t x = 1;
        for (uint32_t e = 0; e < N; ++e) {
            pw[e] = static_cast<uint32_t>(x);
            x = x * r % p;
        }
        uint32_t* o = out + 2 * static_cast<size_t>(j - j0) * N;
        const uint32_t pinv = 0u - pl.pneg[j];
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t w = pw[pl.logN ? bitrev32(i, pl.logN) : 0];
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
                    (static_cast<uint64_t>(pow_mod(r, (base << e) % (2ull * pl.N), p)) << 32) % p);
                z[2 * (rho * pl.logS + e)] = w;
                z[2 * (rho * pl.logS + e) + 1] = shoup_companion(w, p);
            }
        }
        uint32_t* o = om + static_cast<size_t>(j - j0) * (S / 2);
        for (uint32_t b = 0; b < S / 2; ++b) {
            const uint64_t e = pl.logS > 1 ? 2ull * R * bitrev32(b, pl.logS - 1) : 0;
            o[b] = pow_mod(r, e % (2ull * pl.N), p);
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
