// Host-side plan for one point (N, q): the prime set, K, and every constant the
// kernels read. Built once per point, outside any timed region.
//
// The prime and K rule is the reference's (rns_plan.h:138-155): primes
// p < 2^31 with p = 1 (mod max(2N, 2^16)), taken downward from the largest,
// until sum log2 p exceeds 2 * bits(q) + log2 N + 1 + 20.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace pm {

struct HostPlan {
    uint32_t N = 0, logN = 0, W = 0;
    int L = 0, K = 0, qbits = 0;
    std::vector<uint32_t> q;          // L limbs

    std::vector<uint32_t> p, pneg;    // K; pneg = -p^-1 mod 2^32
    std::vector<uint32_t> psi, ipsi;  // K: primitive 2N-th roots of unity and inverses
    std::vector<uint32_t> pow2;       // K x L: 2^(32 l) mod p_j
    // K x L: scale_j 2^(32 l) mod p_j, scale_j the pointwise constant below.
    // Decomposing b with these gives b scale_j mod p_j, so in a product the
    // fused row pass's Montgomery product alone yields a b N^-1 (P/p_j)^-1.
    std::vector<uint32_t> pow2b;
    // Pointwise constant N^-1 (P/p_j)^-1 2^32 mod p_j (the 2^32 undoes the
    // Montgomery product), with its Shoup companion.
    std::vector<uint32_t> scale, scale_s;
    std::vector<double> p_recip;      // K: 1 / p_j
    std::vector<uint32_t> M;          // K x L: P / p_j mod q
    std::vector<uint32_t> negvP;      // (K + 1) x L: -v P mod q
    std::vector<double> mq;           // K: M_j / q
    std::vector<double> nq;           // K + 1: negvP_v / q

    // Geometry of the two-pass NTT: N = R x S, rows of S contiguous elements.
    uint32_t logS = 0, logR = 0;
};

// Throws std::runtime_error with the reference's messages for invalid points
// (rns_plan.h:120-155).
HostPlan make_plan(uint32_t N, uint32_t W, const uint32_t* q, int L);

// Two-pass NTT geometry: rows of S = 2^ceil(logN / 2), columns of R = N / S
// (one pass, R = 1, when N <= 128).
void ntt_geometry(uint32_t logN, uint32_t& logS, uint32_t& logR);

// The prime walk alone: the K primes for (N, bits(q)), or an exception.
std::vector<uint32_t> choose_primes(uint32_t N, int qbits);

// Twiddles w = psi^bitrev_N(i) (forward) or psi^-bitrev_N(i) (inverse), i < count
// (count a power of two <= N), for primes [j0, j1) of the plan, `count` pairs per
// prime: forward as Shoup pairs (w, floor(w 2^32 / p)), inverse as Montgomery
// pairs (w 2^32 mod p, that times p^-1 mod 2^32), which the inverse butterflies
// take for one VALU op less. The first R entries are all the column pass reads.
void build_twiddles(const HostPlan& pl, bool forward, int j0, int j1, uint32_t* out_pairs, uint32_t count);

// Row-pass twiddle factors for primes [j0, j1): the twiddle psi^bitrev_N(idx)
// of row rho (S contiguous elements), stage distance 2^e, block b is
//     Z[rho][e] * Om[b],  Z[rho][e] = psi^(2^e (2 bitrev_R(rho) + 1)),
//                         Om[b]     = psi^(2 R bitrev_{logS-1}(b)),  b < S/2,
// (psi^-1 for the inverse). z_pairs: (R x logS) (Z, Shoup companion) per
// prime; om: S/2 values per prime.
void build_row_twiddles(const HostPlan& pl, bool forward, int j0, int j1, uint32_t* z_pairs, uint32_t* om);

// Balanced base-256 digits: x = sum_i d_i 256^i with d_i in [-128, 127].
// Used for the int8 matrix-core forms.
void balanced_digits(const uint32_t* x, int limbs, int8_t* out, int ndigits);

}  // namespace pm
