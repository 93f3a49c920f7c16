// Butterflies and the register phase shared by the NTT passes (pm_ntt.hip) and
// the fused decompose + first column stages (pm_mfma.hip): word arithmetic
// modulo a prime p < 2^31 on wave64, as device-only inline functions.
//
// Reference counterparts: rns_kernels.cuh:159-174 (the CT / GS butterflies).
#pragma once

#include "pm_arith.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm {
namespace {

__device__ __forceinline__ void ct(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t v = mul_shoup(y, w.x, w.y, p);
    y = sub_mod(x, v, p);
    x = add_mod(x, v, p);
}

// x - y + p lies in (0, 2p), below 2^32, which the Shoup product accepts as is.
__device__ __forceinline__ void gs(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t u = x;
    x = add_mod(u, y, p);
    y = mul_shoup(u - y + p, w.x, w.y, p);
}

// The forward butterfly for any p < 2^31 with residues in [0, 2p): x comes
// down to [0, p); the lazy Shoup remainder v of y w (in [0, 2p)) is formed as
// m = -v - 1 mod 2^32, and max(m, m + p) = -v_r - 1 with v_r = v mod p in
// [0, p) (the max picks m + p exactly when it does not wrap, that is when
// v >= p + 1, and otherwise m, whose v is then p or below; v = p itself gives
// m + p = 2^32 - 1). Then x + v_r = ~M + x and x - v_r + p = x + M + p + 1
// take one instruction each: 10 VALU ops against 12 for the canonical form.
__device__ __forceinline__ void ct_half(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t m = mulhi32(y, w.y) * p - y * w.x - 1u;
    const uint32_t M = max(m, m + p);
    const uint32_t xr = csub(x, p);
    x = ~M + xr;
    y = xr + M + (p + 1u);
}

// The same butterflies with a Montgomery twiddle pair (w.x = w' = w 2^32 mod p,
// canonical, and w.y = w' p^-1 mod 2^32; mont_tw_diff in pm_arith.h): the
// inverse tables hold this form, which saves one VALU op per inverse
// butterfly, and the row pass generates it for 7 VALU ops against 13 for a
// Shoup pair. The product's difference of high words d lies in (-p, p).

// Canonical CT: x, y in [0, p); v = d mod p.
__device__ __forceinline__ void ct_m(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t v = canon_signed(mont_tw_diff(y, w.x, w.y, p), p);
    y = sub_mod(x, v, p);
    x = add_mod(x, v, p);
}

// ct_half's form (p < 2^31, x in [0, 2p)): a = h2 + ~h1 = -d - 1, and
// max(a, a - p) = -v - 1 with v = d mod p (for d >= 0, a >= 2^32 - p and a - p
// does not wrap; for d < 0, a < p and a - p wraps above it). 10 VALU ops, as
// ct_half.
__device__ __forceinline__ void ct_half_m(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t h1 = mulhi32(y, w.x), h2 = mulhi32(y * w.y, p);
    const uint32_t a = h2 + ~h1;
    const uint32_t M = max(a, a - p);
    const uint32_t xr = csub(x, p);
    x = ~M + xr;
    y = xr + M + (p + 1u);
}

// Canonical GS: x, y in [0, p); x - y in (-p, p) as a signed word,
// |(x - y) w'| < p^2.
__device__ __forceinline__ void gs_m(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    const uint32_t r = mont_tw_diff_signed(x - y, w.x, w.y, p);
    x = add_mod(x, y, p);
    y = canon_signed(r, p);
}

// Butterfly forms (kMode): 0 canonical, residues in [0, p); 1 forward with
// residues in [0, 2p) (ct_half), inverse canonical. kMont: the twiddle is a
// Montgomery pair (generated row twiddles, inverse tables), else a Shoup pair.
template <bool kFwd, int kMode, bool kMont>
__device__ __forceinline__ void bfly(uint32_t& x, uint32_t& y, const uint2 w, uint32_t p) {
    if constexpr (kMont) {
        if (kFwd && kMode == 1)
            ct_half_m(x, y, w, p);
        else if (kFwd)
            ct_m(x, y, w, p);
        else
            gs_m(x, y, w, p);
    } else {
        if (kFwd && kMode == 1)
            ct_half(x, y, w, p);
        else if (kFwd)
            ct(x, y, w, p);
        else
            gs(x, y, w, p);
    }
}

template <bool kFwd, bool kTwo, int kMode, class TW>
__device__ __forceinline__ void phase_b(uint32_t (&x)[16], uint32_t (&y)[16], const TW& tw, uint32_t t,
                                        uint32_t p) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int e = kFwd ? 3 - s : s;
        const int d = 1 << e;
#pragma unroll
        for (int blk = 0; blk < 8 / d; ++blk) {
            const uint2 w = tw.b(e, t, blk);
#pragma unroll
            for (int k = blk * 2 * d; k < blk * 2 * d + d; ++k) {
                bfly<kFwd, kMode, TW::kMont>(x[k], x[k + d], w, p);
                if (kTwo) bfly<kFwd, kMode, TW::kMont>(y[k], y[k + d], w, p);
            }
        }
    }
}

__device__ __forceinline__ uint32_t padded(uint32_t u) { return u + (u >> 4); }

// The M / 16 threads of a row are consecutive lanes of one wave, so a row's
// LDS exchange needs only wave-level ordering (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A raw buffer over `bytes` bytes from `base` (gfx9 resource word 3): a lane's
// address is its byte offset plus a wave-uniform offset in a scalar register,
// so strided accesses cost no per-access address arithmetic, and accesses past
// `bytes` read 0 / are dropped by the hardware.
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const void* base, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000)) {}
    __device__ __forceinline__ uint32_t ld(uint32_t voff, uint32_t soff) const {
        return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
    }
    __device__ __forceinline__ void st(uint32_t v, uint32_t voff, uint32_t soff) const {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 0);
    }
    __device__ __forceinline__ uint2 ld2(uint32_t voff, uint32_t soff) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        return make_uint2(v[0], v[1]);
    }
};

}  // namespace
}  // namespace pm
