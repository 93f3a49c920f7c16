// Multi-limb helpers of the plain CRT kernel (pm_device.hip) and the rounding
// steps both CRT kernels share: every loop runs over a compile-time bound (so
// the limbs stay in registers) with the run-time L applied as a guard. The
// reference reduces mod q by a quotient estimate from q's top limbs followed by
// repeated subtraction (rns.cuh:205-221); here the quotient of the whole sum is
// estimated from per-prime FP64 ratios, subtracted inside the carry chain
// (q_times_limb), and the result, in [-q, 2q), is fixed by one signed
// correction (fix_range).
#pragma once

#include "pm_device.h"

namespace pm {

// x (LMAX + 2 limbs, two's complement) holds R = T - Q q with R in [-q, 2q):
// returns R mod q in x, branch-free: R + q if R < 0, R - q if q <= R, else R.
template <int LMAX>
__device__ __forceinline__ void fix_range(uint32_t (&x)[LMAX + 2], const uint32_t* __restrict__ q, int L) {
    const bool neg = static_cast<int32_t>(x[LMAX + 1]) < 0;
    int64_t b = 0;   // borrow of x - q, not stored
#pragma unroll
    for (int i = 0; i < LMAX + 2; ++i) b = (static_cast<int64_t>(x[i]) - (i < L ? q[i] : 0u) + b) >> 32;
    const int64_t sign = neg ? 1 : (b >= 0 ? -1 : 0);   // add q, subtract q, or keep
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < LMAX + 2; ++i) {
        const int64_t s = static_cast<int64_t>(x[i]) + sign * static_cast<int64_t>(i < L ? q[i] : 0u) + c;
        x[i] = static_cast<uint32_t>(s);
        c = s >> 32;
    }
}

// One limb of the subtraction of Qh 2^32 q + Ql q inside a carry chain: the
// products of the previous limbs are carried in a1 (Ql q[i-1]), b1, b2 (Qh q[i-1],
// Qh q[i-2]); returns the 32-bit pieces that land on limb i, summed.
__device__ __forceinline__ int64_t q_times_limb(uint32_t Ql, uint32_t Qh, uint32_t qi, uint64_t& a1, uint64_t& b1,
                                                uint64_t& b2) {
    const uint64_t a = static_cast<uint64_t>(Ql) * qi;
    const uint64_t b = static_cast<uint64_t>(Qh) * qi;
    const int64_t piece = static_cast<int64_t>(static_cast<uint32_t>(a)) + static_cast<int64_t>(a1 >> 32) +
                          static_cast<int64_t>(static_cast<uint32_t>(b1)) + static_cast<int64_t>(b2 >> 32);
    a1 = a;
    b2 = b1;
    b1 = b;
    return piece;
}

// floor(T / q) from its FP64 estimate, clamped to [0, 2^40).
__device__ __forceinline__ uint64_t quotient_estimate(double qf) {
    const double f = floor(qf);
    return f <= 0.0 ? 0ull : static_cast<uint64_t>(f);
}

__device__ __forceinline__ int crt_round(double frac, int K) {
    int v = static_cast<int>(floor(frac + 0.5));
    return v < 0 ? 0 : (v > K ? K : v);
}

}  // namespace pm
