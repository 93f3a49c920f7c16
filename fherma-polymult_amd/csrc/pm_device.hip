// Plain (VALU) kernels for gfx950: limb-to-residue decompose, the two-pass
// negacyclic NTT / inverse NTT with the pointwise product fused into the row
// pass, and the CRT back to limbs mod q.
//
// The NTT is the merged Cooley-Tukey transform with twiddle psi^bitrev_N(i)
// for butterfly group i (reference rns_kernels.cuh:139-174): no separate
// negacyclic twist, bit-reversed order in the NTT domain. Its stages are split
// by butterfly distance t: t >= S runs in the column pass (elements with equal
// index mod S), t < S in the row pass (S contiguous elements). In both passes a
// workgroup owns one tile in LDS for all of its stages. A tile is the set
//     { base + c + u S : c < C, u < U }
// of one prime's row, stored in LDS at u C + c; the butterfly of global
// distance t whose lower element sits at global position k uses twiddle index
//     (N + k) / (2 t),
// which is N/(2t) + k/(2t): the group index of the merged transform. One
// formula therefore serves both passes and both directions.
#include "pm_device.h"

#include "pm_arith.h"
#include "pm_bigint.h"

#include <algorithm>
#include <stdexcept>

namespace pm {
namespace {

// --------------------------------------------------------------------------
// Decompose: res[j][n] = x_n mod p_j with x_n = sum_l limb_l 2^(32 l).
// Each product limb_l * (2^(32 l) mod p_j) is below 2^63 and is summed exactly
// in 64 bits with a wrap counter, then reduced once (rns.cuh:110-128 sums lazy
// Shoup terms instead; the value is the same).

constexpr int kDecCoef = 64;      // coefficients per workgroup, one per lane
constexpr int kDecThreads = 256;  // 4 waves; wave w takes primes w, w + 4, ...

__global__ void __launch_bounds__(kDecThreads)
k_decompose_valu(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res, DevPlan P, uint32_t k0,
                 uint32_t count) {
    extern __shared__ uint32_t dec_tile[];   // kDecCoef rows of L + 1 words (odd stride)
    const int L = P.L;
    const int stride = L + 1;
    const uint32_t n0 = k0 + blockIdx.x * kDecCoef;
    const uint32_t nk = min(static_cast<uint32_t>(kDecCoef), k0 + count - n0);
    const uint32_t* src = limbs + static_cast<size_t>(n0) * L;
    for (uint32_t i = threadIdx.x; i < nk * L; i += blockDim.x) {
        const uint32_t r = i / L;
        dec_tile[r * stride + (i - r * L)] = src[i];
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (lane >= nk) return;
    const uint32_t* x = dec_tile + lane * stride;
    for (int j = wave; j < P.K; j += kDecThreads / 64) {
        const PrimeConst pc = P.pc[j];
        const uint32_t* w = P.pow2 + static_cast<size_t>(j) * L;
        uint64_t acc = 0;
        uint32_t wraps = 0;
        for (int l = 0; l < L; ++l) {
            const uint64_t prod = static_cast<uint64_t>(x[l]) * w[l];
            acc += prod;
            wraps += acc < prod;
        }
        // value = wraps 2^64 + acc; fold to t < 2^64, then t mod p.
        const uint64_t t = static_cast<uint64_t>(static_cast<uint32_t>(acc >> 32)) * pc.c32 +
                           static_cast<uint32_t>(acc) + static_cast<uint64_t>(wraps) * pc.c64;
        const uint32_t hi = mul_shoup(static_cast<uint32_t>(t >> 32), pc.c32, pc.c32s, pc.p);
        const uint32_t lo = mul_shoup(static_cast<uint32_t>(t), 1u, pc.one_s, pc.p);
        res[static_cast<size_t>(j) * P.N + n0 + lane] = add_mod(hi, lo, pc.p);
    }
}

// --------------------------------------------------------------------------
// NTT passes.

constexpr int kNttThreads = 256;

struct TileGeom {
    uint32_t logC, logU;   // tile C x U
    bool col;              // column pass: stages t >= S; row pass: t < S
};

__device__ __forceinline__ uint32_t tile_global(uint32_t i, uint32_t base, const TileGeom& g, uint32_t logS) {
    return base + (i & ((1u << g.logC) - 1)) + ((i >> g.logC) << logS);
}

__device__ __forceinline__ uint32_t tile_base(const TileGeom& g, uint32_t logS) {
    return g.col ? (blockIdx.x << g.logC) : (blockIdx.x << (g.logC + g.logU));
}

// In-tile distances d = 2^e, e in [e_lo, e_hi); global distance 2^logt.
__device__ __forceinline__ void tile_stage_range(const TileGeom& g, uint32_t& e_lo, uint32_t& e_hi) {
    e_lo = g.col ? g.logC : 0;
    e_hi = g.col ? g.logC + g.logU : g.logC;
}

__device__ __forceinline__ uint32_t tile_logt(uint32_t e, const TileGeom& g, uint32_t logS) {
    return g.col ? e - g.logC + logS : e;
}

template <bool kFwd>
__device__ __forceinline__ void tile_stages(uint32_t* tile, const uint2* __restrict__ tw, uint32_t p,
                                            uint32_t base, const TileGeom& g, const DevPlan& P) {
    uint32_t e_lo, e_hi;
    tile_stage_range(g, e_lo, e_hi);
    const uint32_t half = 1u << (g.logC + g.logU - 1);
    for (uint32_t s = 0; s < e_hi - e_lo; ++s) {
        const uint32_t e = kFwd ? e_hi - 1 - s : e_lo + s;
        const uint32_t logt = tile_logt(e, g, P.logS);
        for (uint32_t b = threadIdx.x; b < half; b += blockDim.x) {
            const uint32_t lo = ((b >> e) << (e + 1)) | (b & ((1u << e) - 1));
            const uint32_t hi = lo + (1u << e);
            const uint32_t k = tile_global(lo, base, g, P.logS);
            const uint2 w = tw[(P.N + k) >> (logt + 1)];
            const uint32_t x = tile[lo], y = tile[hi];
            if (kFwd) {
                const uint32_t v = mul_shoup(y, w.x, w.y, p);
                tile[lo] = add_mod(x, v, p);
                tile[hi] = sub_mod(x, v, p);
            } else {
                // inverse tables hold Montgomery pairs; x - y in (-p, p) as a signed word
                tile[lo] = add_mod(x, y, p);
                tile[hi] = canon_signed(mont_tw_diff_signed(x - y, w.x, w.y, p), p);
            }
        }
        __syncthreads();
    }
}

template <bool kFwd>
__global__ void __launch_bounds__(kNttThreads)
k_ntt_pass(uint32_t* __restrict__ X, DevPlan P, TileGeom g) {
    extern __shared__ uint32_t ntt_tile[];
    const int j = blockIdx.y;
    uint32_t* x = X + static_cast<size_t>(j) * P.N;
    const uint2* tw = (kFwd ? P.tw : P.itw) + static_cast<size_t>(j) * P.N;
    const uint32_t p = P.pc[j].p;
    const uint32_t base = tile_base(g, P.logS);
    const uint32_t T = 1u << (g.logC + g.logU);
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) ntt_tile[i] = x[tile_global(i, base, g, P.logS)];
    __syncthreads();
    tile_stages<kFwd>(ntt_tile, tw, p, base, g, P);
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) x[tile_global(i, base, g, P.logS)] = ntt_tile[i];
}

// A * B * N^-1 (P/p_j)^-1: a Montgomery product, then the Shoup product with
// scale = N^-1 (P/p_j)^-1 2^32 (rns_kernels.cuh:178-181).
__device__ __forceinline__ uint32_t pointwise(uint32_t a, uint32_t b, const PrimeConst& pc) {
    return mul_shoup(mont_mul(a, b, pc.p, pc.pneg), pc.scale, pc.scale_s, pc.p);
}

// Row pass fused: forward rows of A and B, pointwise into A, inverse rows of A.
// B must come from the b-side decompose (b_side in pm_device.h).
__global__ void __launch_bounds__(kNttThreads)
k_ntt_rows_fused(uint32_t* __restrict__ A, const uint32_t* __restrict__ B, DevPlan P, TileGeom g) {
    extern __shared__ uint32_t ntt_tile[];
    const uint32_t T = 1u << (g.logC + g.logU);
    uint32_t* ta = ntt_tile;
    uint32_t* tb = ntt_tile + T;
    const int j = blockIdx.y;
    const size_t row = static_cast<size_t>(j) * P.N;
    const PrimeConst pc = P.pc[j];
    const uint32_t base = tile_base(g, P.logS);
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) {
        ta[i] = A[row + base + i];
        tb[i] = B[row + base + i];
    }
    __syncthreads();
    tile_stages<true>(ta, P.tw + row, pc.p, base, g, P);
    tile_stages<true>(tb, P.tw + row, pc.p, base, g, P);
    // B holds b scale_j (b-side decompose): the Montgomery product is the pointwise term
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) ta[i] = mont_mul(ta[i], tb[i], pc.p, pc.pneg);
    __syncthreads();
    tile_stages<false>(ta, P.itw + row, pc.p, base, g, P);
    for (uint32_t i = threadIdx.x; i < T; i += blockDim.x) A[row + base + i] = ta[i];
}

__global__ void k_pointwise(uint32_t* __restrict__ A, const uint32_t* __restrict__ B, DevPlan P) {
    const int j = blockIdx.y;
    const PrimeConst pc = P.pc[j];
    const size_t row = static_cast<size_t>(j) * P.N;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < P.N; n += gridDim.x * blockDim.x) {
        A[row + n] = pointwise(A[row + n], B[row + n], pc);
    }
}

// --------------------------------------------------------------------------
// CRT (VALU): one thread per coefficient n.
//   v  = round(sum_j y_j / p_j), clamped to [0, K]    (FP64; exact thanks to the
//        2^20 margin: the sum lies within 2^-21 of an integer)
//   x  = sum_j y_j M_j + negvP_v                       (LMAX + 2 limbs)
//   c_n = x mod q                                      (quotient estimated from the
//        top limbs in FP64, one below the estimate, then at most two subtractions)
// Reference: rns.cuh:205-238, rns_kernels.cuh:265-364. LMAX is L rounded up to
// a multiple of 8; limbs past L are zero.

template <int LMAX>
__global__ void __launch_bounds__(128)
k_crt_valu(const uint32_t* __restrict__ Y, uint32_t* __restrict__ c, DevPlan P, uint32_t k0, uint32_t count) {
    const uint32_t n = k0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= k0 + count) return;
    const int L = P.L, K = P.K;
    uint64_t col[LMAX + 2];
#pragma unroll
    for (int l = 0; l < LMAX + 2; ++l) col[l] = 0;
    double frac = 0, qf = 0;
    for (int j = 0; j < K; j += 2) {
        const bool two = j + 1 < K;
        const uint32_t y0 = Y[static_cast<size_t>(j) * P.N + n];
        const uint32_t y1 = two ? Y[static_cast<size_t>(j + 1) * P.N + n] : 0u;
        frac = fma(static_cast<double>(y0), P.p_recip[j], frac);
        qf = fma(static_cast<double>(y0), P.mq[j], qf);
        if (two) {
            frac = fma(static_cast<double>(y1), P.p_recip[j + 1], frac);
            qf = fma(static_cast<double>(y1), P.mq[j + 1], qf);
        }
        const uint32_t* m0 = P.M + static_cast<size_t>(j) * L;
        const uint32_t* m1 = two ? m0 + L : m0;
#pragma unroll
        for (int l = 0; l < LMAX; ++l) {
            if (l < L) {
                // two products below 2^63 each: their sum fits 64 bits
                const uint64_t s = static_cast<uint64_t>(y0) * m0[l] + static_cast<uint64_t>(y1) * m1[l];
                col[l] += static_cast<uint32_t>(s);
                col[l + 1] += s >> 32;
            }
        }
    }
    const int v = crt_round(frac, K);
    const uint32_t* nv = P.negvP + static_cast<size_t>(v) * L;
    uint32_t x[LMAX + 2];
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l < LMAX + 2; ++l) {
        const uint64_t s = col[l] + (l < L ? nv[l] : 0u) + carry;
        x[l] = static_cast<uint32_t>(s);
        carry = s >> 32;
    }
    // minus Q q, Q = floor(T / q) from its FP64 estimate (exact or one off)
    const uint64_t Q = quotient_estimate(qf + P.nq[v]);
    const uint32_t Ql = static_cast<uint32_t>(Q), Qh = static_cast<uint32_t>(Q >> 32);
    uint64_t a1 = 0, b1 = 0, b2 = 0;
    int64_t borrow = 0;
#pragma unroll
    for (int l = 0; l < LMAX + 2; ++l) {
        const int64_t s = static_cast<int64_t>(x[l]) - q_times_limb(Ql, Qh, l < L ? P.q[l] : 0u, a1, b1, b2) + borrow;
        x[l] = static_cast<uint32_t>(s);
        borrow = s >> 32;
    }
    fix_range<LMAX>(x, P.q, L);
    uint32_t* out = c + static_cast<size_t>(n) * L;
#pragma unroll
    for (int l = 0; l < LMAX; ++l)
        if (l < L) out[l] = x[l];
}

TileGeom col_geom(const DevPlan& P) {
    TileGeom g;
    g.col = true;
    g.logU = P.logR;
    const uint32_t room = 13 > P.logR ? 13 - P.logR : 0;   // tiles of at most 8192 elements
    g.logC = P.logS < room ? P.logS : room;
    return g;
}

TileGeom row_geom(const DevPlan& P, uint32_t max_log_tile) {
    TileGeom g;
    g.col = false;
    g.logC = P.logS;
    const uint32_t room = max_log_tile > P.logS ? max_log_tile - P.logS : 0;
    g.logU = P.logR < room ? P.logR : room;
    return g;
}

}  // namespace

__global__ void __launch_bounds__(256) k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Read-only pass: per workgroup, the sum mod 2^64 of the 32-bit words of its
// grid-stride share of 16-byte units; a second one-workgroup launch adds the
// partials. One 8-byte write per workgroup: a yardstick for the read counters.
__global__ void __launch_bounds__(256) k_sum16(uint64_t* __restrict__ part, const uint4* __restrict__ src, size_t n) {
    __shared__ uint64_t red[256];
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    uint64_t acc = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4 v = src[i];
        acc += static_cast<uint64_t>(v.x) + v.y + v.z + v.w;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum64(uint64_t* __restrict__ out, const uint64_t* __restrict__ part,
                                               unsigned count) {
    __shared__ uint64_t red[256];
    uint64_t acc = 0;
    for (unsigned i = threadIdx.x; i < count; i += blockDim.x) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

void launch_sum16(uint64_t* partials, uint64_t* out, const void* src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    const unsigned blocks = static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, kSumBlocks)));
    launch(k_sum16, dim3(blocks), dim3(256), 0, st, partials, static_cast<const uint4*>(src), n);
    launch(k_sum64, dim3(1), dim3(256), 0, st, out, static_cast<const uint64_t*>(partials), blocks);
}

void launch_copy16(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (n == 0) return;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 8192));
    launch(k_copy16, dim3(blocks), dim3(256), 0, st, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), n);
}

void launch_decompose_valu(const DevPlan& P, const uint32_t* limbs, uint32_t* res, hipStream_t st) {
    launch_decompose_valu_range(P, limbs, res, 0, P.N, st);
}

void launch_decompose_valu_range(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0,
                                 uint32_t count, hipStream_t st) {
    if (count == 0) return;
    const dim3 grid((count + kDecCoef - 1) / kDecCoef);
    const size_t lds = sizeof(uint32_t) * kDecCoef * (P.L + 1);
    launch(k_decompose_valu, grid, dim3(kDecThreads), lds, st, limbs, res, P, k0, count);
}

void launch_ntt_cols_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    if (P.logR == 0) return;
    if (!P.tw || !P.itw) throw std::logic_error("the generic NTT kernels need the full twiddle tables");
    const TileGeom g = col_geom(P);
    const dim3 grid(1u << (P.logS - g.logC), P.K);
    const size_t lds = sizeof(uint32_t) << (g.logC + g.logU);
    if (forward)
        launch(k_ntt_pass<true>, grid, dim3(kNttThreads), lds, st, X, P, g);
    else
        launch(k_ntt_pass<false>, grid, dim3(kNttThreads), lds, st, X, P, g);
}

void launch_ntt_rows_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    if (!P.tw || !P.itw) throw std::logic_error("the generic NTT kernels need the full twiddle tables");
    const TileGeom g = row_geom(P, 13);
    const dim3 grid(1u << (P.logR - g.logU), P.K);
    const size_t lds = sizeof(uint32_t) << (g.logC + g.logU);
    if (forward)
        launch(k_ntt_pass<true>, grid, dim3(kNttThreads), lds, st, X, P, g);
    else
        launch(k_ntt_pass<false>, grid, dim3(kNttThreads), lds, st, X, P, g);
}

void launch_ntt_rows_fused_generic(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st) {
    if (!P.tw || !P.itw) throw std::logic_error("the generic NTT kernels need the full twiddle tables");
    const TileGeom g = row_geom(P, 12);
    const dim3 grid(1u << (P.logR - g.logU), P.K);
    const size_t lds = 2 * (sizeof(uint32_t) << (g.logC + g.logU));
    launch(k_ntt_rows_fused, grid, dim3(kNttThreads), lds, st, A, B, P, g);
}

void launch_pointwise(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st) {
    const dim3 grid((P.N + 255) / 256 < 64 ? (P.N + 255) / 256 : 64, P.K);
    launch(k_pointwise, grid, dim3(256), 0, st, A, B, P);
}

void launch_crt_valu(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st) {
    launch_crt_valu_range(P, Y, c, 0, P.N, st);
}

void launch_crt_valu_range(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count,
                           hipStream_t st) {
    if (count == 0) return;
    const dim3 grid((count + 127) / 128);
    const int lmax = (P.L + 7) / 8 * 8;
    switch (lmax) {
#define PM_CRT_CASE(LM) \
    case LM: launch(k_crt_valu<LM>, grid, dim3(128), 0, st, Y, c, P, k0, count); break;
        PM_CRT_CASE(8)
        PM_CRT_CASE(16)
        PM_CRT_CASE(24)
        PM_CRT_CASE(32)
        PM_CRT_CASE(40)
        PM_CRT_CASE(48)
        PM_CRT_CASE(56)
        PM_CRT_CASE(64)
#undef PM_CRT_CASE
        default: break;
    }
}

}  // namespace pm
