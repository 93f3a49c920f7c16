/*
 * polymult_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, CPU-only restatement of the reference's exact negacyclic product
 *     c = a * b mod (X^N + 1, q)
 * used by tests/ (as the checker), by __graft_entry__.smoke() (as the checker)
 * and by bench.py's cpu_baseline leg (as the timed CPU port). The product
 * library never links, loads or calls anything in oracle/.
 *
 * It follows the reference algorithm step by step, written for clarity rather
 * than speed (64-bit '%' for every modular product):
 *   plan       rns_plan.h:118-249   (prime walk and K rule 138-155, pneg 178-180,
 *                                    pow2 182-185, psi search 187-194, scale 211-219,
 *                                    M_j 224-236, negvP 238-247)
 *   decompose  rns.cuh:96-128       (x mod p_j = sum_l limb_l * (2^(32 l) mod p_j))
 *   forward    rns.cuh:130-148, rns_kernels.cuh:159-166 (merged Cooley-Tukey, twiddle
 *              psi^bitrev(m + i), bit-reversed output)
 *   pointwise  rns_kernels.cuh:176-181 (A * B * N^-1 * (P/p_j)^-1)
 *   inverse    rns.cuh:150-155, rns_kernels.cuh:168-174 (Gentleman-Sande, psi^-bitrev)
 *   crt        rns.cuh:223-238, rns_kernels.cuh:328-364 (v = round(sum y_j / p_j),
 *              sum_j y_j M_j + negvP_v, reduced mod q)
 * The reduction mod q (reference rns.cuh:205-221, a floating-point quotient
 * estimate) is restated as exact binary long division: x mod q is defined by the
 * mathematics, so the two agree whenever both are correct.
 *
 * Also here: the counter-based SplitMix64 input generator shared with the
 * fixtures (tests/golden/make_golden.py restates it in pure Python) and an
 * O(N^2) multi-limb schoolbook product for tiny N.
 *
 * Pinning (tests/test_oracle.py): gmpy2 Kronecker digests in tests/golden/,
 * the reference's own solve_cpu.cpp built into oracle/_ref, closed-form
 * known answers (zero, identity, wrap, all-(q-1), alternating), the NTT checked
 * against its definition and against the convolution theorem, decompose and CRT
 * checked against Python big integers.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define PMO_MAX_L 64
#define PMO_MAX_K 160
#define PMO_MAX_LOGN 22
#define PMO_MARGIN_BITS 20

/* ------------------------------------------------------------------------- */
/* Word arithmetic mod p < 2^31 (reference rns.cuh:46-93, restated with '%'). */

static uint32_t pow_mod(uint64_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1;
    b %= p;
    for (; e; e >>= 1, b = b * b % p)
        if (e & 1) r = r * b % p;
    return (uint32_t)r;
}

static uint32_t mul_mod(uint32_t a, uint32_t b, uint32_t p) {
    return (uint32_t)((uint64_t)a * b % p);
}

/* Deterministic Miller-Rabin, bases 2, 3, 5, 7 (exact below 3.2e9), as
 * rns_plan.h:63-83 does. */
static int is_prime_u32(uint32_t n) {
    static const uint32_t bases[4] = {2, 3, 5, 7};
    if (n < 2) return 0;
    for (int i = 0; i < 4; ++i)
        if (n % bases[i] == 0) return n == bases[i];
    uint32_t d = n - 1;
    int r = 0;
    while (!(d & 1)) d >>= 1, ++r;
    for (int i = 0; i < 4; ++i) {
        uint64_t x = pow_mod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int k = 1; k < r && composite; ++k) {
            x = x * x % n;
            if (x == n - 1) composite = 0;
        }
        if (composite) return 0;
    }
    return 1;
}

static uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* ------------------------------------------------------------------------- */
/* Multi-limb helpers: little-endian u32 limbs. */

static int bit_length(const uint32_t* x, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (x[i]) return 32 * i + 32 - __builtin_clz(x[i]);
    return 0;
}

/* Compare x (n limbs) with y << (32*0 + s bits) ... done bitwise by long division
 * below; this compares x with y, both n limbs. */
static int cmp_n(const uint32_t* x, const uint32_t* y, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (x[i] != y[i]) return x[i] > y[i] ? 1 : -1;
    return 0;
}

static void sub_n(uint32_t* x, const uint32_t* y, int n) {
    uint64_t borrow = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t d = (uint64_t)x[i] - y[i] - borrow;
        x[i] = (uint32_t)d;
        borrow = (d >> 63) & 1;
    }
}

/* y = q << s (s bits), into n limbs (caller guarantees it fits). */
static void shl_into(uint32_t* y, int n, const uint32_t* q, int L, int s) {
    int w = s / 32, b = s % 32;
    memset(y, 0, sizeof(uint32_t) * (size_t)n);
    for (int i = 0; i < L; ++i) {
        uint64_t v = (uint64_t)q[i] << b;
        if (i + w < n) y[i + w] |= (uint32_t)v;
        if (i + w + 1 < n) y[i + w + 1] |= (uint32_t)(v >> 32);
    }
}

/* x mod q in place by binary long division: x has n limbs (n <= 2 * PMO_MAX_L + 4),
 * q has L limbs with a non-zero top limb. Result in x[0..L). */
static void reduce_mod_q(uint32_t* x, int n, const uint32_t* q, int L) {
    uint32_t y[2 * PMO_MAX_L + 8];
    const int qb = bit_length(q, L);
    int xb = bit_length(x, n);
    for (int s = xb - qb; s >= 0; --s) {
        shl_into(y, n, q, L, s);
        if (cmp_n(x, y, n) >= 0) sub_n(x, y, n);
    }
}

/* x = x * m mod q for x < q (L limbs). */
static void mul_small_mod(uint32_t* x, uint32_t m, const uint32_t* q, int L) {
    uint32_t t[PMO_MAX_L + 1];
    uint64_t carry = 0;
    for (int l = 0; l < L; ++l) {
        uint64_t v = (uint64_t)x[l] * m + carry;
        t[l] = (uint32_t)v;
        carry = v >> 32;
    }
    t[L] = (uint32_t)carry;
    reduce_mod_q(t, L + 1, q, L);
    memcpy(x, t, sizeof(uint32_t) * (size_t)L);
}

/* ------------------------------------------------------------------------- */
/* The plan (rns_plan.h:19-50, 118-249). */

typedef struct pmo_plan {
    uint32_t N, logN;
    int L, K, qbits;
    uint32_t q[PMO_MAX_L];
    uint32_t p[PMO_MAX_K];
    uint32_t pneg[PMO_MAX_K];
    uint32_t psi[PMO_MAX_K], ipsi[PMO_MAX_K];
    uint32_t scale[PMO_MAX_K];        /* N^-1 (P/p_j)^-1 mod p_j (no Montgomery factor) */
    uint32_t* pow2;                   /* K x L: 2^(32 l) mod p_j */
    uint32_t* tw;                     /* K x N: psi^bitrev(i) */
    uint32_t* itw;                    /* K x N: psi^-bitrev(i) */
    uint32_t* M;                      /* K x L: P / p_j mod q */
    uint32_t* negvP;                  /* (K + 1) x L: -v P mod q */
} pmo_plan;

void pmo_plan_free(pmo_plan* pl) {
    if (!pl) return;
    free(pl->pow2);
    free(pl->tw);
    free(pl->itw);
    free(pl->M);
    free(pl->negvP);
    free(pl);
}

static void set_err(char* err, size_t errlen, const char* msg) {
    if (err && errlen) {
        strncpy(err, msg, errlen - 1);
        err[errlen - 1] = 0;
    }
}

/* The prime set and K for (N, qbits): rns_plan.h:138-155 (primes p < 2^31).
 * Returns K, or -1 if there are not enough primes, -2 if K exceeds `cap`. */
int pmo_choose_primes(uint32_t N, int qbits, uint32_t* out, int cap) {
    uint32_t logN = 0;
    while ((1u << logN) < N) ++logN;
    const double need_bits = 2.0 * qbits + logN + 1 + PMO_MARGIN_BITS;
    double have_bits = 0;
    const uint32_t step = 2 * N > (1u << 16) ? 2 * N : (1u << 16);
    int K = 0;
    for (uint32_t c = ((1u << 31) - 1) / step * step + 1; have_bits <= need_bits; c -= step) {
        if (c < step) return -1;
        if (!is_prime_u32(c)) continue;
        if (K >= cap) return -2;
        out[K++] = c;
        have_bits += log2((double)c);
    }
    return K;
}

pmo_plan* pmo_plan_new(uint32_t N, const uint32_t* q, int L, char* err, size_t errlen) {
    if (N < 2 || (N & (N - 1)) || N > (1u << PMO_MAX_LOGN)) {
        set_err(err, errlen, "N must be a power of two in [2, 2^22]");
        return NULL;
    }
    if (L < 1 || L > PMO_MAX_L) {
        set_err(err, errlen, "L must be in [1, 64]");
        return NULL;
    }
    pmo_plan* pl = (pmo_plan*)calloc(1, sizeof(pmo_plan));
    pl->N = N;
    while ((1u << pl->logN) < N) ++pl->logN;
    pl->L = L;
    memcpy(pl->q, q, sizeof(uint32_t) * (size_t)L);
    pl->qbits = bit_length(pl->q, L);
    if (pl->qbits <= 32 * (L - 1)) {
        set_err(err, errlen, "q has fewer limbs than L");
        pmo_plan_free(pl);
        return NULL;
    }
    int K = pmo_choose_primes(N, pl->qbits, pl->p, PMO_MAX_K);
    if (K == -1) {
        set_err(err, errlen, "not enough primes p < 2^31 with p = 1 (mod step)");
        pmo_plan_free(pl);
        return NULL;
    }
    if (K < 0) {
        set_err(err, errlen, "K exceeds the maximum of 160");
        pmo_plan_free(pl);
        return NULL;
    }
    pl->K = K;
    pl->pow2 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K * L);
    pl->tw = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K * N);
    pl->itw = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)K * N);
    pl->M = (uint32_t*)calloc((size_t)K * L, sizeof(uint32_t));
    pl->negvP = (uint32_t*)calloc((size_t)(K + 1) * L, sizeof(uint32_t));

    for (int j = 0; j < K; ++j) {
        const uint32_t p = pl->p[j];
        uint32_t inv = 1; /* p^-1 mod 2^32 by Newton (rns_plan.h:178-180) */
        for (int i = 0; i < 5; ++i) inv *= 2 - p * inv;
        pl->pneg[j] = 0u - inv;
        for (int l = 0; l < L; ++l) pl->pow2[(size_t)j * L + l] = pow_mod(2, 32ull * (uint64_t)l, p);
        /* psi: the first g^((p-1)/2N), g = 2, 3, ..., of order exactly 2N
         * (rns_plan.h:187-191). */
        uint32_t psi = 0;
        for (uint32_t g = 2; !psi; ++g) {
            const uint32_t cand = pow_mod(g, (p - 1) / (2ull * N), p);
            if (pow_mod(cand, N, p) == p - 1) psi = cand;
        }
        pl->psi[j] = psi;
        pl->ipsi[j] = pow_mod(psi, p - 2, p);
        /* N^-1 (P/p_j)^-1 mod p_j (rns_plan.h:211-219, without the 2^32 that
         * undoes the reference's Montgomery product). */
        uint64_t Pj = 1;
        for (int i = 0; i < K; ++i)
            if (i != j) Pj = Pj * (pl->p[i] % p) % p;
        const uint64_t n_inv = pow_mod(N, p - 2, p);
        const uint64_t Pj_inv = pow_mod(Pj, p - 2, p);
        pl->scale[j] = (uint32_t)(n_inv * Pj_inv % p);
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < K; ++j) {
        const uint32_t p = pl->p[j];
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t e = bitrev(i, pl->logN);
            pl->tw[(size_t)j * N + i] = pow_mod(pl->psi[j], e, p);
            pl->itw[(size_t)j * N + i] = pow_mod(pl->ipsi[j], e, p);
        }
    }
    /* M_j = prod_{i != j} p_i mod q, and P mod q (rns_plan.h:224-236). */
    uint32_t Pq[PMO_MAX_L];
    memset(Pq, 0, sizeof(Pq));
    Pq[0] = 1;
    for (int j = 0; j < K; ++j) {
        uint32_t m[PMO_MAX_L];
        memset(m, 0, sizeof(m));
        m[0] = 1;
        for (int i = 0; i < K; ++i)
            if (i != j) mul_small_mod(m, pl->p[i], pl->q, L);
        memcpy(pl->M + (size_t)j * L, m, sizeof(uint32_t) * (size_t)L);
        mul_small_mod(Pq, pl->p[j], pl->q, L);
    }
    /* negvP_v = (q - v P mod q) mod q, v = 0..K (rns_plan.h:238-247). */
    for (int v = 1; v <= K; ++v) {
        uint32_t t[PMO_MAX_L];
        memcpy(t, Pq, sizeof(uint32_t) * (size_t)L);
        mul_small_mod(t, (uint32_t)v, pl->q, L);
        if (bit_length(t, L) == 0) continue;
        uint32_t neg[PMO_MAX_L];
        memcpy(neg, pl->q, sizeof(uint32_t) * (size_t)L);
        sub_n(neg, t, L);
        memcpy(pl->negvP + (size_t)v * L, neg, sizeof(uint32_t) * (size_t)L);
    }
    return pl;
}

int pmo_plan_K(const pmo_plan* pl) { return pl->K; }
uint32_t pmo_plan_logN(const pmo_plan* pl) { return pl->logN; }
const uint32_t* pmo_plan_primes(const pmo_plan* pl) { return pl->p; }
const uint32_t* pmo_plan_psi(const pmo_plan* pl) { return pl->psi; }
const uint32_t* pmo_plan_M(const pmo_plan* pl) { return pl->M; }
const uint32_t* pmo_plan_negvP(const pmo_plan* pl) { return pl->negvP; }
const uint32_t* pmo_plan_scale(const pmo_plan* pl) { return pl->scale; }

/* ------------------------------------------------------------------------- */
/* The five steps. Residues are K x N, prime-major. */

/* rns.cuh:110-128: x mod p = sum_l limb_l * (2^(32 l) mod p). */
void pmo_decompose(const pmo_plan* pl, const uint32_t* limbs, uint32_t* res) {
    const uint32_t N = pl->N;
    const int L = pl->L, K = pl->K;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < K; ++j) {
        const uint32_t p = pl->p[j];
        const uint32_t* w = pl->pow2 + (size_t)j * L;
        for (uint32_t n = 0; n < N; ++n) {
            uint64_t acc = 0;
            const uint32_t* x = limbs + (size_t)n * L;
            for (int l = 0; l < L; ++l) acc = (acc + (uint64_t)(x[l] % p) * w[l]) % p;
            res[(size_t)j * N + n] = (uint32_t)acc;
        }
    }
}

/* Forward negacyclic NTT of one row, merged Cooley-Tukey: for t = N/2 .. 1,
 * group i of the stage with m = N / (2t) groups uses psi^bitrev(m + i)
 * (rns_kernels.cuh:139-157 reduce to this global indexing). */
static void ntt_fwd_row(uint32_t* x, uint32_t N, const uint32_t* tw, uint32_t p) {
    for (uint32_t m = 1, t = N / 2; m < N; m <<= 1, t >>= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t w = tw[m + i];
            for (uint32_t k = 2 * i * t; k < 2 * i * t + t; ++k) {
                const uint32_t u = x[k], v = mul_mod(x[k + t], w, p);
                x[k] = (u + v) % p;
                x[k + t] = (u + p - v) % p;
            }
        }
    }
}

/* Inverse, Gentleman-Sande: for t = 1 .. N/2, (x, y) -> (x + y, (x - y) w). */
static void ntt_inv_row(uint32_t* x, uint32_t N, const uint32_t* itw, uint32_t p) {
    for (uint32_t m = N / 2, t = 1; m >= 1; m >>= 1, t <<= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t w = itw[m + i];
            for (uint32_t k = 2 * i * t; k < 2 * i * t + t; ++k) {
                const uint32_t u = x[k], v = x[k + t];
                x[k] = (u + v) % p;
                x[k + t] = mul_mod((u + p - v) % p, w, p);
            }
        }
    }
}

void pmo_ntt_forward(const pmo_plan* pl, uint32_t* res) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < pl->K; ++j)
        ntt_fwd_row(res + (size_t)j * pl->N, pl->N, pl->tw + (size_t)j * pl->N, pl->p[j]);
}

void pmo_ntt_inverse(const pmo_plan* pl, uint32_t* res) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < pl->K; ++j)
        ntt_inv_row(res + (size_t)j * pl->N, pl->N, pl->itw + (size_t)j * pl->N, pl->p[j]);
}

/* rns_kernels.cuh:178-181: A * B * N^-1 * (P/p_j)^-1 into A. */
void pmo_pointwise(const pmo_plan* pl, uint32_t* A, const uint32_t* B) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < pl->K; ++j) {
        const uint32_t p = pl->p[j], s = pl->scale[j];
        uint32_t* a = A + (size_t)j * pl->N;
        const uint32_t* b = B + (size_t)j * pl->N;
        for (uint32_t n = 0; n < pl->N; ++n) a[n] = mul_mod(mul_mod(a[n], b[n], p), s, p);
    }
}

/* v for one coefficient: round(sum_j y_j / p_j), clamped to [0, K]
 * (rns_kernels.cuh:333-337). Also returns the distance to the nearest integer,
 * which the 2^20 margin keeps below 2^-21. */
static int crt_v(const pmo_plan* pl, const uint32_t* Y, uint32_t n, double* dist) {
    double frac = 0;
    for (int j = 0; j < pl->K; ++j) frac += (double)Y[(size_t)j * pl->N + n] / (double)pl->p[j];
    double r = floor(frac + 0.5);
    if (dist) *dist = fabs(frac - r);
    int v = (int)r;
    return v < 0 ? 0 : (v > pl->K ? pl->K : v);
}

/* rns_kernels.cuh:265-364: c = (sum_j y_j M_j + negvP_v) mod q. Returns the largest
 * distance of sum_j y_j / p_j from an integer seen over the N coefficients. */
double pmo_crt(const pmo_plan* pl, const uint32_t* Y, uint32_t* c) {
    const uint32_t N = pl->N;
    const int L = pl->L, K = pl->K;
    double worst = 0;
#pragma omp parallel for schedule(static) reduction(max : worst)
    for (uint32_t n = 0; n < N; ++n) {
        double dist;
        const int v = crt_v(pl, Y, n, &dist);
        if (dist > worst) worst = dist;
        uint64_t col[PMO_MAX_L + 3];
        for (int l = 0; l < L + 3; ++l) col[l] = l < L ? pl->negvP[(size_t)v * L + l] : 0;
        for (int j = 0; j < K; ++j) {
            const uint64_t y = Y[(size_t)j * N + n];
            const uint32_t* m = pl->M + (size_t)j * L;
            for (int l = 0; l < L; ++l) {
                const uint64_t prod = y * m[l];
                col[l] += (uint32_t)prod;
                col[l + 1] += prod >> 32;
            }
        }
        uint32_t x[PMO_MAX_L + 3];
        uint64_t carry = 0;
        for (int l = 0; l < L + 3; ++l) {
            const uint64_t s = col[l] + carry;
            x[l] = (uint32_t)s;
            carry = s >> 32;
        }
        reduce_mod_q(x, L + 3, pl->q, L);
        memcpy(c + (size_t)n * L, x, sizeof(uint32_t) * (size_t)L);
    }
    return worst;
}

/* The whole product, in the reference's phase order (solve_cpu.cpp:221-243). */
double pmo_multiply(const pmo_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    const size_t KN = (size_t)pl->K * pl->N;
    uint32_t* A = (uint32_t*)malloc(sizeof(uint32_t) * KN);
    uint32_t* B = (uint32_t*)malloc(sizeof(uint32_t) * KN);
    pmo_decompose(pl, a, A);
    pmo_decompose(pl, b, B);
    pmo_ntt_forward(pl, A);
    pmo_ntt_forward(pl, B);
    pmo_pointwise(pl, A, B);
    pmo_ntt_inverse(pl, A);
    double worst = pmo_crt(pl, A, c);
    free(A);
    free(B);
    return worst;
}

void pmo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int pmo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Seeded inputs: counter-based SplitMix64, restated in make_golden.py.
 *   key(seed, stream) = mix(seed * G + stream * H)
 *   word(i)           = mix(key + (i + 1) * G) >> 32
 * q (stream 0): L words, masked to W bits, bits W-1 and 0 set.
 * operand (stream s): coefficient n, attempt t (< 64), limb l uses word
 * ((n * 64 + t) * L + l); the top limb is masked to W bits; the first attempt
 * below q is kept (each attempt succeeds with probability q / 2^W >= 1/2). */

#define PMO_G 0x9E3779B97F4A7C15ull
#define PMO_H 0xD1B54A32D192ED03ull

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t stream_key(uint64_t seed, uint64_t stream) { return mix64(seed * PMO_G + stream * PMO_H); }

static uint32_t stream_word(uint64_t key, uint64_t i) { return (uint32_t)(mix64(key + (i + 1) * PMO_G) >> 32); }

static uint32_t top_mask(uint32_t W) {
    const uint32_t r = W % 32;
    return r ? (1u << r) - 1 : 0xFFFFFFFFu;
}

void pmo_gen_q(uint64_t seed, uint32_t W, uint32_t* q) {
    const uint32_t L = (W + 31) / 32;
    const uint64_t key = stream_key(seed, 0);
    for (uint32_t l = 0; l < L; ++l) q[l] = stream_word(key, l);
    q[L - 1] &= top_mask(W);
    q[L - 1] |= 1u << ((W - 1) % 32);
    q[0] |= 1u;
}

/* Returns 0, or -1 if some coefficient needed more than 64 attempts. */
int pmo_gen_operand(uint64_t seed, uint32_t stream, const uint32_t* q, uint32_t W, uint32_t N, uint32_t* out) {
    const uint32_t L = (W + 31) / 32;
    const uint64_t key = stream_key(seed, stream);
    const uint32_t mask = top_mask(W);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (uint32_t n = 0; n < N; ++n) {
        uint32_t* x = out + (size_t)n * L;
        int ok = 0;
        for (uint32_t t = 0; t < 64 && !ok; ++t) {
            for (uint32_t l = 0; l < L; ++l) x[l] = stream_word(key, ((uint64_t)n * 64 + t) * L + l);
            x[L - 1] &= mask;
            ok = cmp_n(x, q, (int)L) < 0;
        }
        if (!ok) bad = 1;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* O(N^2 L^2) schoolbook negacyclic product mod q, for tiny N: every product
 * a_i b_j is reduced mod q and added to or subtracted from c_(i+j mod N). */
void pmo_schoolbook(uint32_t N, const uint32_t* q, int L, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    memset(c, 0, sizeof(uint32_t) * (size_t)N * L);
    uint32_t prod[2 * PMO_MAX_L + 2];
    for (uint32_t i = 0; i < N; ++i) {
        for (uint32_t j = 0; j < N; ++j) {
            memset(prod, 0, sizeof(prod));
            const uint32_t* x = a + (size_t)i * L;
            const uint32_t* y = b + (size_t)j * L;
            for (int u = 0; u < L; ++u) {
                uint64_t carry = 0;
                for (int v = 0; v < L; ++v) {
                    const uint64_t t = (uint64_t)x[u] * y[v] + prod[u + v] + carry;
                    prod[u + v] = (uint32_t)t;
                    carry = t >> 32;
                }
                int k = u + L;
                while (carry) {
                    const uint64_t t = (uint64_t)prod[k] + carry;
                    prod[k++] = (uint32_t)t;
                    carry = t >> 32;
                }
            }
            reduce_mod_q(prod, 2 * L + 1, q, L);
            const uint32_t k = (i + j) % N;
            uint32_t* z = c + (size_t)k * L;
            uint32_t t[PMO_MAX_L + 1];
            if (i + j < N) { /* z = (z + prod) mod q */
                uint64_t carry = 0;
                for (int l = 0; l < L; ++l) {
                    const uint64_t s = (uint64_t)z[l] + prod[l] + carry;
                    t[l] = (uint32_t)s;
                    carry = s >> 32;
                }
                t[L] = (uint32_t)carry;
                reduce_mod_q(t, L + 1, q, L);
            } else { /* z = (z + q - prod) mod q */
                uint64_t carry = 0;
                for (int l = 0; l < L; ++l) {
                    const uint64_t s = (uint64_t)z[l] + q[l] + carry;
                    t[l] = (uint32_t)s;
                    carry = s >> 32;
                }
                t[L] = (uint32_t)carry;
                uint32_t pp[PMO_MAX_L + 1];
                memcpy(pp, prod, sizeof(uint32_t) * (size_t)L);
                pp[L] = 0;
                sub_n(t, pp, L + 1);
                reduce_mod_q(t, L + 1, q, L);
            }
            memcpy(z, t, sizeof(uint32_t) * (size_t)L);
        }
    }
}
