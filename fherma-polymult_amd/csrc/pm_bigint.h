// Multi-limb helpers for the CRT kernels: reduction of an (LMAX + 2)-limb
// value mod q, with every loop over a compile-time bound (so the limbs stay in
// registers) and the run-time L applied as a guard. Reference counterpart:
// rns.cuh:157-221 (geq, sub_in_place, submul, reduce_wide).
#pragma once

#include "pm_device.h"

namespace pm {

template <int LMAX>
__device__ __forceinline__ bool geq_q(const uint32_t (&x)[LMAX + 2], const uint32_t* __restrict__ q, int L) {
    bool decided = false, ge = true;
#pragma unroll
    for (int i = LMAX + 1; i >= 0; --i) {
        const uint32_t qi = i < L ? q[i] : 0u;
        if (!decided && x[i] != qi) {
            decided = true;
            ge = x[i] > qi;
        }
    }
    return ge;
}

template <int LMAX>
__device__ __forceinline__ void sub_q(uint32_t (&x)[LMAX + 2], const uint32_t* __restrict__ q, int L) {
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < LMAX + 2; ++i) {
        const uint64_t d = static_cast<uint64_t>(x[i]) - (i < L ? q[i] : 0u) - borrow;
        x[i] = static_cast<uint32_t>(d);
        borrow = static_cast<uint32_t>(d >> 63);
    }
}

// x -= e q 2^(32 shift), for a result known to be non-negative.
template <int LMAX>
__device__ __forceinline__ void sub_mul_q(uint32_t (&x)[LMAX + 2], const uint32_t* __restrict__ q, int L,
                                          uint32_t e, int shift) {
    uint64_t carry = 0;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < LMAX + 2; ++i) {
        if (i < shift) continue;
        const int k = i - shift;
        const uint64_t prod = (k < L ? static_cast<uint64_t>(e) * q[k] : 0ull) + carry;
        carry = prod >> 32;
        const uint64_t d = static_cast<uint64_t>(x[i]) - static_cast<uint32_t>(prod) - borrow;
        x[i] = static_cast<uint32_t>(d);
        borrow = static_cast<uint32_t>(d >> 63);
    }
}

template <int LMAX>
__device__ __forceinline__ void reduce_mod_q(uint32_t (&x)[LMAX + 2], const DevPlan& P, const uint32_t* q) {
    double xd = 0;
#pragma unroll
    for (int i = LMAX + 1; i >= 0; --i) {
        if (i >= P.q_drop) xd = xd * 4294967296.0 + x[i];
    }
    const double f = floor(xd / P.q_top);
    const uint64_t e = f >= 1.0 ? static_cast<uint64_t>(f) - 1 : 0;
    if (e >> 32) sub_mul_q<LMAX>(x, q, P.L, static_cast<uint32_t>(e >> 32), 1);
    if (static_cast<uint32_t>(e)) sub_mul_q<LMAX>(x, q, P.L, static_cast<uint32_t>(e), 0);
    for (int r = 0; r < 3 && geq_q<LMAX>(x, q, P.L); ++r) sub_q<LMAX>(x, q, P.L);
}

// The same with q read from the plan's device copy.
template <int LMAX>
__device__ __forceinline__ void reduce_mod_q(uint32_t (&x)[LMAX + 2], const DevPlan& P) {
    reduce_mod_q<LMAX>(x, P, P.q);
}

__device__ __forceinline__ int crt_round(double frac, int K) {
    int v = static_cast<int>(floor(frac + 0.5));
    return v < 0 ? 0 : (v > K ? K : v);
}

}  // namespace pm
