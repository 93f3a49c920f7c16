// Word arithmetic modulo a prime p < 2^31, shared by the host plan and the
// gfx950 kernels.
//
// With the reference's primes (just below 2^31) a 32-bit word holds residues
// up to 2p but not 4p, so sums are reduced with one unsigned min each; the NTT
// kernels' butterfly forms built on these (residues in [0, 2p) forward, and
// lazily reduced ones for primes below 2^30) are in pm_ntt.hip.
//
// Reference counterparts: rns.cuh:46-93 (mulhi, add_mod, sub_mod, mul_shoup,
// mont_mul, bitrev_bits).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PM_HD __host__ __device__ __forceinline__
#else
#define PM_HD inline
#endif

namespace pm {

constexpr int kMaxL = 64;        // limbs per coefficient (W <= 2048), rns.cuh:37
constexpr int kMaxK = 160;       // primes, rns.cuh:38
constexpr int kMaxLogN = 22;     // rns.cuh:39
constexpr int kMarginBits = 20;  // sum_j y_j / p_j lies within 2^-21 of an integer, rns.cuh:41

PM_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
#endif
}

// Reduce r in [0, 2p) to [0, p): if r < p then r - p wraps above r.
PM_HD uint32_t csub(uint32_t r, uint32_t p) {
    const uint32_t t = r - p;
    return t < r ? t : r;
}

PM_HD uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) { return csub(a + b, p); }

PM_HD uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b;
    const uint32_t e = d + p;
    return e < d ? e : d;   // a < b: d wrapped high and d + p is the answer, below d
}

// a * w mod p, w < p constant with ws = floor(w 2^32 / p); any a < 2^32.
PM_HD uint32_t mul_shoup(uint32_t a, uint32_t w, uint32_t ws, uint32_t p) {
    return csub(a * w - mulhi32(a, ws) * p, p);
}

// The same, left in [0, 2p).
PM_HD uint32_t mul_shoup_lazy(uint32_t a, uint32_t w, uint32_t ws, uint32_t p) {
    return a * w - mulhi32(a, ws) * p;
}

// a * b * 2^-32 mod p, left in [0, 2p), for a b < 2^32 p; pneg = -p^-1 mod 2^32.
PM_HD uint32_t mont_mul_lazy(uint32_t a, uint32_t b, uint32_t p, uint32_t pneg) {
    const uint32_t lo = a * b;
    const uint32_t hi = mulhi32(a, b);
    const uint32_t m = lo * pneg;
    // (a b + m p) / 2^32: the low words cancel, with a carry out unless lo == 0.
    return hi + mulhi32(m, p) + (lo != 0u);
}

// The same reduced to [0, p), for a b < 2^32 p (a, b < p, or a < p and b < 2p).
PM_HD uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t pneg) {
    return csub(mont_mul_lazy(a, b, p, pneg), p);
}

// High word of the signed 64-bit product.
PM_HD int32_t mulhi_i32(int32_t a, int32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __mulhi(a, b);
#else
    return static_cast<int32_t>((static_cast<int64_t>(a) * b) >> 32);
#endif
}

// Twiddle products with a Montgomery pair (wm = w 2^32 mod p, canonical, and
// wq = wm p^-1 mod 2^32). With m = y wq mod 2^32, m p and y wm agree in their
// low words, so (y wm - m p) / 2^32 = hi(y wm) - hi(m p) exactly, congruent to
// y w mod p. For any y < 2^32, y wm < 2^32 p puts both high words in [0, p) and
// the result, as a signed word, in (-p, p).
PM_HD uint32_t mont_tw_diff(uint32_t y, uint32_t wm, uint32_t wq, uint32_t p) {
    return mulhi32(y, wm) - mulhi32(y * wq, p);
}

// The same for a signed d (the inverse butterflies' x - y) with |d wm| < 2^31 p,
// from the signed high words: again in (-p, p).
PM_HD uint32_t mont_tw_diff_signed(uint32_t d, uint32_t wm, uint32_t wq, uint32_t p) {
    const int32_t hi_dw = mulhi_i32(static_cast<int32_t>(d), static_cast<int32_t>(wm));
    const int32_t hi_mp = mulhi_i32(static_cast<int32_t>(d * wq), static_cast<int32_t>(p));
    return static_cast<uint32_t>(hi_dw - hi_mp);
}

// r in (-p, p) as a signed word, brought to [0, p): r + p falls below r (it
// wraps) exactly when r is negative.
PM_HD uint32_t canon_signed(uint32_t r, uint32_t p) {
    const uint32_t s = r + p;
    return s < r ? s : r;
}

PM_HD uint32_t shoup_companion(uint32_t w, uint32_t p) {
    return static_cast<uint32_t>((static_cast<uint64_t>(w) << 32) / p);
}

PM_HD uint32_t bitrev32(uint32_t x, uint32_t bits) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __brev(x) >> (32 - bits);
#else
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
#endif
}

}  // namespace pm
