// Exact int8 matrix-core (MFMA) forms of decompose and CRT for gfx950.
//
// Both steps are contractions of a big-integer operand against a per-plan basis,
// which on the VALU cost N L K 32x32-bit multiply-adds each. Here they become
// int8 x int8 -> int32 products on v_mfma_i32_16x16x64_i8, with every partial
// sum bounded so that the int32 accumulators cannot overflow; the result is
// then reduced exactly (mod p_j, or mod q), so the output is bit-identical to
// the plain form's.
//
// Operand bytes: an unsigned byte u becomes the int8 digit u - 128 (u ^ 0x80),
// and the missing 128 * sum(...) is added back as a precomputed constant. Basis
// values become balanced base-256 digits in [-128, 127], built on the host.
//
// Decompose (reference rns.cuh:110-128: x mod p_j = sum_l limb_l 2^(32 l) mod p_j):
//   x_n = sum_t u_{n,t} 256^t over its 4L bytes; T_{j,t} = 256^t 2^32 mod p_j
//   (Montgomery form), centred into (-p/2, p/2] and split into 4 balanced
//   digits e_{j,t,b}:
//     G_b[j][n] = sum_t e_{j,t,b} (u_{n,t} - 128)        |G_b| <= 4L 2^14 <= 2^22
//     x_n 2^32 mod p_j = (sum_b 256^b G_b + 128 sum_t T_{j,t}) mod p_j
//   MFMA rows are (prime, digit) pairs, columns coefficients, K-dim bytes t. The
//   four G_b of one (j, n) land in one lane; the bias term and a positive offset
//   enter as the accumulators' initial values, the four sums are combined into
//   one exact positive 64-bit integer and reduced by one Montgomery step, which
//   also takes the factor 2^32 away: integer arithmetic only (FP64 appears in
//   the CRT's quotient rounding alone).
//
// CRT (reference rns_kernels.cuh:265-364: x = sum_j y_j M_j + negvP_v mod q):
//   y_{j,n} = sum_a u_{j,n,a} 256^a (4 bytes), M_j = sum_s m_{j,s} 256^s (4L + 1
//   balanced digits). Then 256^a M_j has digits m_{j,s-a}, so
//     G[s][n] = sum_{(j,a)} m_{j,s-a} (u_{j,n,a} - 128)   |G| <= 4K 2^14 < 2^24
//     sum_j y_j M_j = sum_s 256^s G[s][n] + 0x80808080 sum_j M_j   (exactly)
//   MFMA rows are byte positions s, columns coefficients, K-dim (prime, byte).
//   v = round(sum_j y_j / p_j) and the FP64 quotient estimate Q are formed
//   from sums accumulated while the y bytes are loaded; one more K-dim group,
//   the quotient group, holds the balanced digits of Q and v against table
//   columns -q, -q 2^32 and -P mod q, so the contraction itself yields
//   R = sum_j y_j M_j + v (-P mod q) - Q q. A lane's four rows are the four
//   bytes of one 32-bit limb, so it folds them into one signed 64-bit limb
//   partial; the same kernel then runs the carry chain and reduces mod q
//   (k_crt_mfma below).
//
// MFMA operand layout (16x16x64 i8): lane l supplies row (A) / column (B)
// l % 16 and the 16 K-elements 16 (l / 16) + [0, 16) of each 64-wide K step;
// the result D[i][j] sits in lane 16 (i / 4) + j, element i % 4. The K order
// inside a lane only has to agree between A and B, which it does by
// construction.
#include "pm_arith.h"
#include "pm_bfly.h"
#include "pm_bigint.h"
#include "pm_device.h"
#include "pm_mfma_host.h"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;          // waves per workgroup
constexpr int kTiles = 4;          // 16-coefficient tiles per wave: 64 coefficients
constexpr int kDecMaxKSteps = 4;   // 4L <= 256 bytes
constexpr int kCrtMaxKSteps = 11;  // 16 (ceil(K / 4) + 1) <= 704: the primes and the quotient group

__device__ __forceinline__ v4i mfma16(const v4i& a, const v4i& b, const v4i& c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// The digits of the words past L in each of a tile's 64 LDS rows: byte 0,
// stored as 0x80 (the bias of u - 128).
template <int KT>
__device__ __forceinline__ void pad_rows(uint32_t* xs, int L) {
    constexpr int Ls = 16 * KT + 4;
    for (uint32_t r = threadIdx.x; r < 16u * kTiles; r += blockDim.x)
        for (int w = L; w < Ls; ++w) xs[r * Ls + w] = 0x80808080u;
}

// Decompose, one workgroup per 64 coefficients (4 tiles of 16). Their limbs are
// copied to LDS with coalesced loads (rows of 16 KT + 4 words, zero beyond L,
// stored as int8 digits u ^ 0x80), every wave takes its B fragments from there,
// and the four waves split the prime groups. A fragments of the next group are
// loaded before the current group's MFMAs. kLazyOut: residues left in [0, 2p).
template <int KT, bool kLazyOut>
__device__ __forceinline__ void decompose_tile(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res,
                                               const DevPlan& P, uint32_t k0, uint32_t count) {
    extern __shared__ uint32_t xs[];
    constexpr int Ls = 16 * KT + 4;
    const int L = P.L;
    // per-prime constants: the accumulators' initial values, p, -p^-1 mod 2^32
    int2* kinit = reinterpret_cast<int2*>(xs + 16 * kTiles * Ls);
    uint32_t* kp = reinterpret_cast<uint32_t*>(kinit + P.K);
    uint32_t* kpneg = kp + P.K;
    for (int j = threadIdx.x; j < P.K; j += blockDim.x) {
        kinit[j] = P.dec_init[j];
        kp[j] = P.pc[j].p;
        kpneg[j] = P.pc[j].pneg;
    }
    const uint32_t nb = k0 + blockIdx.x * (16 * kTiles);
    const uint32_t nk = min(static_cast<uint32_t>(16 * kTiles), k0 + count - nb);
    const uint32_t* src = limbs + static_cast<size_t>(nb) * L;
    constexpr uint32_t kSlots = 16u * kTiles * Ls;
    if (nk == 16u * kTiles && (L & 3) == 0 && L >= 8) {
        // A full tile of whole 16-byte units per row: its 64 L words are
        // contiguous and 16-byte aligned (nb is a multiple of 64), and each unit
        // lands in one LDS row (Ls is a multiple of 4). All loads in flight.
        const uint32_t q4 = static_cast<uint32_t>(L) / 4, units = 16u * kTiles * q4;
        const uint4* src4 = reinterpret_cast<const uint4*>(src);
        uint4 v[KT];
#pragma unroll
        for (int r = 0; r < KT; ++r) v[r] = src4[min(threadIdx.x + r * blockDim.x, units - 1)];
#pragma unroll
        for (int r = 0; r < KT; ++r) {
            const uint32_t i4 = threadIdx.x + r * blockDim.x;
            if (i4 < units) {
                const uint32_t row = mulhi32(i4, P.dec_magic), c4 = i4 - row * q4;
                *reinterpret_cast<uint4*>(xs + row * Ls + 4 * c4) =
                    make_uint4(v[r].x ^ 0x80808080u, v[r].y ^ 0x80808080u, v[r].z ^ 0x80808080u, v[r].w ^ 0x80808080u);
            }
        }
        pad_rows<KT>(xs, L);
    } else if (nk == 16u * kTiles) {
        // A full tile, L not a multiple of 4: at most KT 16-byte loads per thread.
        const uint32_t words = 16u * kTiles * static_cast<uint32_t>(L);
        const uint4* src4 = reinterpret_cast<const uint4*>(src);
        uint4 v[KT];
#pragma unroll
        for (int r = 0; r < KT; ++r) {
            const uint32_t i4 = threadIdx.x + r * blockDim.x;
            v[r] = src4[min(i4, words / 4 - 1)];
        }
#pragma unroll
        for (int r = 0; r < KT; ++r) {
            const uint32_t i4 = threadIdx.x + r * blockDim.x;
            if (4 * i4 < words) {
                const uint32_t w4[4] = {v[r].x, v[r].y, v[r].z, v[r].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    // i / L by the plan's magic multiplier (exact for i < 2^12)
                    const uint32_t i = 4 * i4 + e, row = L == 1 ? i : mulhi32(i, P.dec_magic);
                    xs[row * Ls + (i - row * L)] = w4[e] ^ 0x80808080u;
                }
            }
        }
        pad_rows<KT>(xs, L);
    } else {
        for (uint32_t i = threadIdx.x; i < kSlots; i += blockDim.x) {
            const uint32_t r = i / Ls, w = i - r * Ls;
            const bool in = r < nk && w < static_cast<uint32_t>(L);
            const uint32_t v = src[in ? r * L + w : 0];
            xs[i] = (in ? v : 0u) ^ 0x80808080u;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    v4i bf[kTiles][KT];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int ks = 0; ks < KT; ++ks)
            bf[t][ks] = *reinterpret_cast<const v4i*>(xs + (16 * t + cl) * Ls + 16 * ks + 4 * g);

    const v4i* dig = reinterpret_cast<const v4i*>(P.dec_digits);
    const int groups = (P.K + 3) >> 2;
    v4i an[KT];
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) an[ks] = dig[(min(wave, groups - 1) * KT + ks) * 64 + lane];
    // 2^16, opaque to the compiler: R below is one v_mad_u64_u32 instead of
    // sign extensions and 64-bit shifts
    uint32_t k16 = 1u << 16;
    asm volatile("" : "+s"(k16));
    for (int jg = wave; jg < groups; jg += kWaves) {
        v4i a[KT];
#pragma unroll
        for (int ks = 0; ks < KT; ++ks) a[ks] = an[ks];
        const int jn = min(jg + kWaves, groups - 1);
#pragma unroll
        for (int ks = 0; ks < KT; ++ks) an[ks] = dig[(jn * KT + ks) * 64 + lane];
        // Lane (g, cl) holds rows 4g .. 4g + 3 = prime 4 jg + g, digits b = 0..3;
        // digits 0 and 2 start from the prime's offsets (lanes past K: any
        // prime's, their sums are dropped)
        const int j = 4 * jg + g;
        const int2 init = kinit[min(j, P.K - 1)];
        v4i acc[kTiles];
#pragma unroll
        for (int t = 0; t < kTiles; ++t) acc[t] = v4i{init.x, 0, init.y, 0};
#pragma unroll
        for (int ks = 0; ks < KT; ++ks)
#pragma unroll
            for (int t = 0; t < kTiles; ++t) acc[t] = mfma16(a[ks], bf[t][ks], acc[t]);
        if (j < P.K) {
            const uint32_t p = kp[j], pneg = kpneg[j];
            uint32_t* out = res + static_cast<size_t>(j) * P.N + nb;
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                // lo = G_0 + 256 G_1 fits int32 (|lo| < 2^30 + 2^23); hi, with the
                // offset, is in [2^23, 2^31 + 2^25) as an unsigned word; R = 2^16 hi + lo
                // lies in (0, 2^48) and is congruent to the exact sum. One Montgomery
                // step: t = (R + m p) / 2^32 < p + 2^16, below 2p.
                const int lo = acc[t][0] + acc[t][1] * 256;
                const uint32_t hi = static_cast<uint32_t>(acc[t][2]) + static_cast<uint32_t>(acc[t][3]) * 256u;
                const uint64_t R = static_cast<uint64_t>(hi) * k16 + static_cast<uint64_t>(static_cast<int64_t>(lo));
                const uint32_t m = static_cast<uint32_t>(R) * pneg;
                const uint32_t v = static_cast<uint32_t>((static_cast<uint64_t>(m) * p + R) >> 32);
                if (static_cast<uint32_t>(16 * t + cl) < nk) out[16 * t + cl] = kLazyOut ? v : min(v, v - p);
            }
        }
    }
}

template <int KT, bool kLazyOut>
__global__ void __launch_bounds__(64 * kWaves)
k_decompose_mfma(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res, DevPlan P, uint32_t k0,
                 uint32_t count) {
    decompose_tile<KT, kLazyOut>(limbs, res, P, k0, count);
}

// A batch: blockIdx.z = 2 (product) + (0: a into A, 1: b into B); every
// product's L needs KT K-steps.
template <int KT, bool kLazyOut>
__global__ void __launch_bounds__(64 * kWaves) k_decompose_mfma_batch(DevPlan P, BatchArgs ba) {
    const int i = blockIdx.z >> 1;
    const bool second = (blockIdx.z & 1) != 0;   // b: scaled by the pointwise constant
    const auto d = as_const(ba.d[i]);
    DevPlan Q = P;
    Q.K = d->K;
    Q.L = d->L;
    Q.dec_magic = d->dec_magic;
    Q.dec_digits = second ? d->dec_digits_b : d->dec_digits;
    Q.dec_init = second ? d->dec_init_b : d->dec_init;
    decompose_tile<KT, kLazyOut>(second ? ba.b[i] : ba.a[i], second ? ba.B[i] : ba.A[i], Q, 0, P.N);
}

// Decompose fused with the first forward column stages (DevPlan::fused_cols;
// N = 65536 or 32768, R1 = R / 16 = 16 or 8): coefficient n = s + N1 r with
// N1 = N / R1, r < R1; the stages of distance N/2 .. N1 are R1-point transforms
// over r for each inner position s, with twiddle index (R1 + r) >> (e + 1) for
// the element r of the stage of distance N1 2^e (the column pass's form with
// R1 rows: the first R1 pairs of the prime's forward table). A workgroup of
// R1 / 2 waves takes 32 consecutive s and all R1 r (32 R1 coefficients): wave
// w holds the B fragments of the four tiles r = 4 (w mod R1/4) .. + 3 at
// s = s0 + 16 (w div R1/4) + [0, 16) in registers (loaded straight from the
// limbs, no LDS) and loops over the prime groups [g_begin, g_end). After each
// group the residues go to an LDS exchange buffer [prime of the batch][r][s];
// after every four groups (16 primes) one barrier, and the threads run the
// 16 x 32 (prime, s) columns through the stages in registers (one column per
// thread for R1 = 16, two for R1 = 8) and store them, two primes x 32 s per
// wave and step (128-byte segments). Two exchange buffers alternate, so one
// barrier per batch keeps every read of a buffer ahead of its next write. The
// decompose's [0, 2p) residues feed ct_half as they are (kMode 1).
constexpr int kFcCols = 32;   // inner positions per workgroup

template <int R1>
struct FcShape {
    static constexpr int kWaves = R1 / 2;                 // four tiles per wave
    static constexpr int kRq = R1 / 4;                    // waves per 16-s half
    static constexpr int kStride = R1 * kFcCols + 16;     // exchange words per prime: 16 mod 64 banks apart
    static constexpr int kSteps = 512 / (64 * kWaves);    // transform columns per thread
};

// LDS: two exchange buffers, two sets of the batch's 16 x 16 twiddle pairs,
// the per-prime constants
template <int R1>
__host__ __device__ constexpr uint32_t fc_lds_bytes(int K) {
    return 4u * 2u * 16u * FcShape<R1>::kStride + 2u * 16u * 16u * 8u + 16u * static_cast<uint32_t>(K);
}

// The R1-point forward stages of one column, twiddles from the prime's first
// pairs (Shoup pairs) staged in LDS.
template <int R1, int kMode>
__device__ __forceinline__ void outer_stages(uint32_t (&x)[R1], const uint2* tw, uint32_t p) {
    constexpr int LOG = R1 == 16 ? 4 : 3;
#pragma unroll
    for (int e = LOG - 1; e >= 0; --e) {
        const int d = 1 << e;
#pragma unroll
        for (int blk = 0; blk < R1 / (2 * d); ++blk) {
            const uint2 w = tw[(R1 >> (e + 1)) + blk];
#pragma unroll
            for (int k = blk * 2 * d; k < blk * 2 * d + d; ++k) bfly<true, kMode, false>(x[k], x[k + d], w, p);
        }
    }
}

template <int R1, int KT, int kMode>
__device__ __forceinline__ void decompose_cols_block(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res,
                                                     const DevPlan& P, int g_begin, int g_end) {
    using F = FcShape<R1>;
    extern __shared__ uint32_t xs[];
    const int K = P.K, L = P.L;
    uint2* twl = reinterpret_cast<uint2*>(xs + 2 * 16 * F::kStride);   // [2][16 primes][16 pairs]
    int2* kinit = reinterpret_cast<int2*>(twl + 2 * 16 * 16);
    uint32_t* kp = reinterpret_cast<uint32_t*>(kinit + K);
    uint32_t* kpneg = kp + K;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        kinit[j] = P.dec_init[j];
        kp[j] = P.pc[j].p;
        kpneg[j] = P.pc[j].pneg;
    }
    const uint32_t N1 = P.N / R1;
    const uint32_t s0 = blockIdx.x * kFcCols;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int sh = wave / F::kRq, rq = wave % F::kRq;
    // B fragments: tile t is r = 4 rq + t; lane (g, cl) holds words 16 ks + 4 g
    // .. + 3 of coefficient s0 + 16 sh + cl + N1 r as int8 digits u ^ 0x80,
    // words past L as 0 (digit 0x80)
    v4i bf[kTiles][KT];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const uint32_t n = s0 + 16 * sh + cl + N1 * (4 * rq + t);
        const uint32_t* src = limbs + static_cast<size_t>(n) * L;
#pragma unroll
        for (int ks = 0; ks < KT; ++ks) {
            const int w0 = 16 * ks + 4 * g;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if ((L & 3) == 0) {
                // 16-byte aligned: n L and w0 are multiples of 4, and w0 < L
                // puts all four words below L
                if (w0 < L) v = *reinterpret_cast<const uint4*>(src + w0);
            } else {
                v.x = w0 < L ? src[w0] : 0u;
                v.y = w0 + 1 < L ? src[w0 + 1] : 0u;
                v.z = w0 + 2 < L ? src[w0 + 2] : 0u;
                v.w = w0 + 3 < L ? src[w0 + 3] : 0u;
            }
            bf[t][ks] = v4i{static_cast<int>(v.x ^ 0x80808080u), static_cast<int>(v.y ^ 0x80808080u),
                            static_cast<int>(v.z ^ 0x80808080u), static_cast<int>(v.w ^ 0x80808080u)};
        }
    }
    const v4i* dig = reinterpret_cast<const v4i*>(P.dec_digits);
    const int groups = (K + 3) >> 2;
    v4i an[KT];
#pragma unroll
    for (int ks = 0; ks < KT; ++ks) an[ks] = dig[(min(g_begin, groups - 1) * KT + ks) * 64 + lane];
    uint32_t k16 = 1u << 16;
    asm volatile("" : "+s"(k16));
    __syncthreads();   // the per-prime constants
    int buf = 0;
    for (int jb = g_begin; jb < g_end; jb += 4, buf ^= 1) {
        uint32_t* xe = xs + buf * 16 * F::kStride;
        uint2* twb = twl + buf * 256;
        // the batch's twiddles: thread i < 256 copies pair i % 16 of prime 4 jb + i / 16
        // (a prime past K copies the last prime's, never read)
        if (threadIdx.x < 256)
            twb[threadIdx.x] = P.ctw[(static_cast<size_t>(min(4 * jb + static_cast<int>(threadIdx.x >> 4), K - 1)) << P.logR) +
                                     (threadIdx.x & 15)];
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const int jg = jb + q;
            if (jg >= g_end) break;
            const int j = 4 * jg + g;
            const int2 init = kinit[min(j, K - 1)];
            v4i acc[kTiles];
#pragma unroll
            for (int t = 0; t < kTiles; ++t) acc[t] = v4i{init.x, 0, init.y, 0};
            // the next group's A fragments load while this group's MFMAs and
            // epilogue run (each K-step's fragment is reloaded once its MFMAs
            // are issued)
            const int jn = min(jg + 1, groups - 1);
#pragma unroll
            for (int ks = 0; ks < KT; ++ks) {
#pragma unroll
                for (int t = 0; t < kTiles; ++t) acc[t] = mfma16(an[ks], bf[t][ks], acc[t]);
                an[ks] = dig[(jn * KT + ks) * 64 + lane];
            }
            const uint32_t p = kp[min(j, K - 1)], pneg = kpneg[min(j, K - 1)];
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                // the fold of decompose_tile: R in (0, 2^48), one Montgomery step
                const int lo = acc[t][0] + acc[t][1] * 256;
                const uint32_t hi = static_cast<uint32_t>(acc[t][2]) + static_cast<uint32_t>(acc[t][3]) * 256u;
                const uint64_t Rw = static_cast<uint64_t>(hi) * k16 + static_cast<uint64_t>(static_cast<int64_t>(lo));
                const uint32_t m = static_cast<uint32_t>(Rw) * pneg;
                const uint32_t v = static_cast<uint32_t>((static_cast<uint64_t>(m) * p + Rw) >> 32);
                xe[(4 * q + g) * F::kStride + (4 * rq + t) * kFcCols + 16 * sh + cl] = kMode == 1 ? v : min(v, v - p);
            }
        }
        __syncthreads();
        // the columns: step it takes column it (64 kWaves) + thread; a wave's two
        // primes sit 2 apart (their exchange rows then 32 banks apart), 32 s each
#pragma unroll
        for (int it = 0; it < F::kSteps; ++it) {
            const int idx = it * 64 * F::kWaves + static_cast<int>(threadIdx.x);
            const int tw_ = idx >> 6, th = (idx >> 5) & 1;
            const int tpl = 4 * (tw_ >> 1) + (tw_ & 1) + 2 * th;
            const uint32_t tsl = idx & 31;
            const int j = 4 * jb + tpl;
            if (j < 4 * g_end && j < K) {
                uint32_t x[R1];
#pragma unroll
                for (int r = 0; r < R1; ++r) x[r] = xe[tpl * F::kStride + r * kFcCols + tsl];
                outer_stages<R1, kMode>(x, twb + 16 * tpl, kp[j]);
                uint32_t* out = res + static_cast<size_t>(j) * P.N + s0 + tsl;
#pragma unroll
                for (int r = 0; r < R1; ++r) out[N1 * r] = x[r];
            }
        }
    }
}

// A batch: blockIdx.z = 2 (product) + (0: a into A, 1: b into B), or, for
// side 0 / 1, blockIdx.z = product and that side only; blockIdx.y = the range
// of prime groups [y gps, (y + 1) gps) (gps a multiple of 4).
template <int R1, int KT, int kMode>
__global__ void __launch_bounds__(32 * R1, 4) k_decompose_cols_batch(DevPlan P, BatchArgs ba, int side, int gps) {
    const bool both = side < 0;
    const int i = both ? blockIdx.z >> 1 : blockIdx.z;
    const bool second = both ? (blockIdx.z & 1) != 0 : side == 1;
    const auto d = as_const(ba.d[i]);
    const int groups = (d->K + 3) >> 2;
    const int g0 = static_cast<int>(blockIdx.y) * gps;
    if (g0 >= groups) return;
    DevPlan Q = P;
    Q.K = d->K;
    Q.L = d->L;
    Q.dec_digits = second ? d->dec_digits_b : d->dec_digits;
    Q.dec_init = second ? d->dec_init_b : d->dec_init;
    decompose_cols_block<R1, KT, kMode>(second ? ba.b[i] : ba.a[i], second ? ba.B[i] : ba.A[i], Q, g0,
                                        min(groups, g0 + gps));
}

// CRT, one workgroup (four waves) per 64 coefficients, in one kernel:
//   1. all threads: y[j][nb .. nb + 64) for every prime, byte-transposed per
//      group of four primes (word a of group J holds byte a of y_4J .. y_4J+3),
//      into LDS as int8 digits, with partial sums of y_j / p_j and y_j M_j / q;
//      the M digit table is copied to LDS alongside (once per workgroup). Each
//      wave then forms v = round(sum y_j / p_j) and the quotient estimate
//      Q = floor(sum y_j M_j / q + v (-P mod q) / q) per coefficient and
//      writes their digits as the quotient group's B word.
//   2. MFMA: wave w owns the limb tiles st = w, w + 4, w + 8, w + 12 (wave 0
//      also st = 16) for all four 16-coefficient tiles, a 4 x 4 block of
//      accumulators, so every A and B fragment read from LDS feeds four MFMAs.
//      Each lane folds its four rows (the bytes of one limb) into a signed
//      64-bit limb partial; after a barrier they go to LDS as [limb][coefficient].
//   3. R = sum_l partial_l 2^(32 l) + offset correction (the partials include
//      v (-P mod q) - Q q) is congruent to the product mod q and lies in
//      [-q, 2q) (Q is exact or one off). The L + 2 limbs are cut into four segments, one per wave, lane =
//      coefficient: each wave runs its segment's carry chain with carry-in 0,
//      then the carries between segments are resolved from each segment's
//      carry-out, low 64 bits and all-zero / all-ones flags (a carry-in can
//      move a segment's carry-out by at most one). The sign of R and R >= q
//      (compared segment-wise from the top) pick R + q, R - q or R; that
//      addition is resolved across segments the same way.
//   4. the wave's limbs pass through LDS so the stores are whole lines.
//
// K-dim order: K-step ks, lane group g, fragment word a, byte r is prime
// 4 (4 ks + g) + r, byte a of y; the A word for row s is then word J = 4 ks + g
// of digit-table row s - a.
constexpr int kTabRows = 16 * 17 + 3;   // digit rows -3 .. 271 of the transposed table
constexpr int kTabLoads = 12;           // 16-byte loads per thread for the table: 44 x 275 words
constexpr int kSegMax = 17;             // limbs per segment: 4 x 17 >= L + 2 for L <= 64

// Digit table words, column-major: word (J, R) at J kTabRows + R holds byte R - 3
// of the balanced digits of M_4J .. M_4J+3 (zero outside 0 .. 4L and for primes
// past K); column ceil(K / 4) those of -q, -q 2^32, -P mod q and 0 (the
// quotient group); 4 KS columns in all, the ones that KS K-steps read; whole
// 16-byte units.
__host__ __device__ constexpr uint32_t crt_tab_words(int KS) {
    return 4u * static_cast<uint32_t>(KS) * kTabRows;
}

// LDS words, three regions:
//   R0  y digits [4 KS][64] as 16-byte words (phases 1-2), then the limb
//       partials (int64, [68][64], rows past 4 ST zero), then the output stage
//       (64 x kStage);
//   R1  the digit table, loaded once and kept for every tile of the workgroup;
//   aux FP64 partial sums (1024 words), reused in phase 3 for the segment
//       exports and comparisons.
// Every bound the tile loop tests is a compile-time one (KS) or a fixed size:
// uniform runtime bounds became lane masks that the compiler hoisted out of
// the loop and spilled.
constexpr int kPartRows = 4 * kSegMax;   // 68 limbs: partial rows and phase 3's segments
constexpr int kStage = kPartRows + 1;    // output stage row stride (words)
__host__ __device__ constexpr uint32_t crt_r0_words(int KS) {
    return static_cast<uint32_t>(1024 * KS) > static_cast<uint32_t>(2 * 64 * kPartRows)
               ? static_cast<uint32_t>(1024 * KS)
               : static_cast<uint32_t>(2 * 64 * kPartRows);
}

__host__ __device__ constexpr uint32_t crt_lds_words(int KS) {
    return crt_r0_words(KS) + crt_tab_words(KS) + 1024;
}

// d[a] = byte a of w0, w1, w2, w3 (in that byte order), by byte permutes:
// v_perm_b32 picks bytes 0-3 from its second operand and 4-7 from its first.
__device__ __forceinline__ void transpose_bytes(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t (&d)[4]) {
    const uint32_t lo01 = __builtin_amdgcn_perm(w1, w0, 0x05010400u);   // w0.b0 w1.b0 w0.b1 w1.b1
    const uint32_t hi01 = __builtin_amdgcn_perm(w1, w0, 0x07030602u);   // w0.b2 w1.b2 w0.b3 w1.b3
    const uint32_t lo23 = __builtin_amdgcn_perm(w3, w2, 0x05010400u);
    const uint32_t hi23 = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
    d[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    d[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    d[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    d[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// a0 + 2^8 a1 + 2^16 a2 + 2^24 a3 as a signed 64-bit value. The factors come in
// as opaque scalars (k8 = 2^8, k16 = 2^16, k24 = 2^24), so each term is one
// signed 32 x 32 + 64-bit multiply-add (v_mad_i64_i32) instead of a sign
// extension, a 64-bit shift and a 64-bit add.
__device__ __forceinline__ long long fold_limb(const v4i& a, int k8, int k16, int k24) {
    long long x = static_cast<long long>(a[0]);
    x += static_cast<long long>(a[1]) * k8;
    x += static_cast<long long>(a[2]) * k16;
    x += static_cast<long long>(a[3]) * k24;
    return x;
}

// A segment's low 64 bits, as far as a carry-in |C| < 2^30 can tell: lo itself
// below 2^30, lo - 2^64 above 2^64 - 2^30, else kFar.
constexpr int kFar = -0x7fffffff - 1;
__device__ __forceinline__ int lo_summary(uint64_t lo) {
    const long long s = static_cast<long long>(lo);
    return (lo < (1ull << 30)) ? static_cast<int>(lo) : (s >= -(1ll << 30) && s < 0 ? static_cast<int>(s) : kFar);
}

// Carry out of a segment of kSegMax limbs whose own chain (carry-in 0) ended
// with carry-out cout, low-64-bit summary lo_s and upper limbs all zero (uz) /
// all ones (uo), when carry-in C (|C| < 2^30) is added instead:
// cout + floor((X + C) / 2^(32 kSegMax)).
__device__ __forceinline__ int seg_carry(int C, int cout, int lo_s, bool uz, bool uo) {
    int adj = 0;
    if (C > 0 && uo && lo_s != kFar && lo_s < 0 && lo_s + C >= 0) adj = 1;
    if (C < 0 && uz && lo_s != kFar && lo_s >= 0 && lo_s < -C) adj = -1;
    return cout + adj;
}

// A pointer into the constant address space (scalar loads for uniform
// addresses) whose value the compiler cannot see through: loads from it stay
// where they are written instead of being hoisted out of a loop.
template <class T>
using kconst = const T __attribute__((address_space(4)))*;
template <class T>
__device__ __forceinline__ kconst<T> laundered(const T* p) {
    kconst<T> k = (kconst<T>)p;
    asm volatile("" : "+s"(k));
    return k;
}

// An LDS word pointer whose value the compiler cannot see through (it stays one
// register, and constant offsets from it fold into the DS instructions).
using lds_u32 = const uint32_t __attribute__((address_space(3)))*;
__device__ __forceinline__ lds_u32 lds_opaque(const uint32_t* p) {
    uint32_t a = static_cast<uint32_t>((uintptr_t)(lds_u32)p);
    asm volatile("" : "+v"(a));
    return (lds_u32)(uintptr_t)a;
}

template <int KS>
__device__ __forceinline__ void crt_tiles(const uint32_t* __restrict__ Y, uint32_t* __restrict__ c, const DevPlan& P,
                                          uint32_t k0, uint32_t count, uint32_t tile0, uint32_t tstep) {
    extern __shared__ uint32_t smem[];
    const int K = P.K, L = P.L;
    constexpr uint32_t r0 = crt_r0_words(KS), r1 = crt_tab_words(KS);
    uint4* yq = reinterpret_cast<uint4*>(smem);                                // [4 KS][64]
    long long* part = reinterpret_cast<long long*>(smem);                      // after phase 2: [limb][64]
    uint32_t* stage = smem;                                                    // after phase 3: 64 x kStage
    uint32_t* tab = smem + r0;                                                 // [KJ][kTabRows]
    double* fr = reinterpret_cast<double*>(tab + r1);                          // 4 x 64
    double* fq = fr + 4 * 64;                                                  // 4 x 64
    // phase 3, over fr / fq (dead by then): segment exports, 4 x 64 each
    int* ex_c = reinterpret_cast<int*>(fr);          // carry-out
    int* ex_s = ex_c + 256;                          // low-64-bit summary
    uint32_t* ex_f = reinterpret_cast<uint32_t*>(ex_c + 512);   // all-zero / all-ones flags
    int* cmpw = ex_c + 768;                          // comparison with q
    const uint32_t ntiles = (count + 63) / 64;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: scalar limb indices

    // 1. the digit table, once per workgroup, and the first tile's y words
    {
        const uint4* tab4 = reinterpret_cast<const uint4*>(P.crt_tab);
        constexpr uint32_t tab4n = crt_tab_words(KS) / 4;
        uint4 tv[kTabLoads];
#pragma unroll
        for (int r = 0; r < kTabLoads; ++r) tv[r] = tab4[min(threadIdx.x + r * blockDim.x, tab4n - 1)];
        // past the table's end every thread rewrites its last entry with the same value
        uint4* tab4s = reinterpret_cast<uint4*>(tab);
#pragma unroll
        for (int r = 0; r < kTabLoads; ++r) tab4s[min(threadIdx.x + r * blockDim.x, tab4n - 1)] = tv[r];
    }
    // prime groups J = wave + 4 i, i < KS: primes 4 J + r; all loads of a tile in
    // flight at once, through a buffer over the K x N residues: row j at the
    // scalar offset 4 j N, the lane's column at 4 col; rows past K read 0
    // (4 K N < 2^31: the launchers check it).
    constexpr int kRows = 4 * KS;
    uint32_t y[kRows];
    const __amdgpu_buffer_rsrc_t ybuf = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(Y), 0, static_cast<int>(4u * static_cast<uint32_t>(K) * P.N), 0x00020000);
    auto load_y = [&](uint32_t tile) {
        const uint32_t nb = k0 + tile * 64, nk = min(64u, k0 + count - nb);
        const uint32_t voff = 4u * (nb + min(static_cast<uint32_t>(lane), nk - 1));
        // the row stride is reloaded per tile: loop-invariant row offsets hoisted
        // out of the tile loop would not fit in scalar registers
        uint32_t stride = 4u * P.N;
        asm volatile("" : "+s"(stride));
        uint32_t soff = 4u * static_cast<uint32_t>(wave) * stride;
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            y[i] = __builtin_amdgcn_raw_buffer_load_b32(ybuf, voff, soff, 0);
            soff += (i % 4 == 3) ? 13u * stride : stride;   // next prime, or the next group J + 4
        }
    };
    if (tile0 < ntiles) load_y(tile0);
    // Tiles tile0, tile0 + tstep, ...: the next tile's y words are loaded into
    // registers while this one runs its MFMA and carry phases.
    for (uint32_t tile = tile0; tile < ntiles; tile += tstep) {
    const uint32_t nb = k0 + tile * 64;
    const uint32_t nk = min(64u, k0 + count - nb);
    {
        // Lanes past nk (a last, partial tile) hold copies of coefficient
        // nb + nk - 1 (load_y clamps the column); their results are not stored.
        // The constants are loaded per tile (scalar loads), not hoisted out of
        // the tile loop: 2 x 4 KS doubles would not fit in scalar registers.
        const kconst<double> fc = laundered(P.crt_fc);
        double f = 0, f2 = 0;
#pragma unroll
        for (int i = 0; i < KS; ++i) {
            const int J = wave + 4 * i;
            uint32_t w[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // y rows past K read 0; the constants (160 entries) are zero past K
                const int j = 4 * J + r;
                w[r] = y[4 * i + r];
                f = fma(static_cast<double>(w[r]), fc[j], f);
                f2 = fma(static_cast<double>(w[r]), fc[4 * 4 * kCrtMaxKSteps + j], f2);
            }
            uint32_t d[4];
            transpose_bytes(w[0], w[1], w[2], w[3], d);
            // B word k is byte 3 - k (the A word k is digit row s - 3 + k); the
            // table is zero for primes past K, so their digits add nothing
            yq[J * 64 + lane] = make_uint4(d[3] ^ 0x80808080u, d[2] ^ 0x80808080u, d[1] ^ 0x80808080u,
                                           d[0] ^ 0x80808080u);
        }
        fr[wave * 64 + lane] = f;
        fq[wave * 64 + lane] = f2;
    }
    __syncthreads();
    {
        // v and Q per coefficient, in every wave (no barrier before phase 2: all
        // waves write the same quotient-group word, each before reading it)
        const double frac = fr[lane] + fr[64 + lane] + fr[128 + lane] + fr[192 + lane];
        const int v = crt_round(frac, K);
        // -P mod q is negvP row 1; v (-P mod q) / q is v nq[1]
        const uint64_t Q = quotient_estimate(fq[lane] + fq[64 + lane] + fq[128 + lane] + fq[192 + lane] +
                                             static_cast<double>(v) * P.nq[1]);
        // the quotient group: "prime" 0 holds Q's balanced digits 0-3 (table
        // column -q), 1 digits 4-5 (-q 2^32), 2 v's two digits (-P mod q), as
        // signed bytes without the bias
        long long xq = static_cast<long long>(Q), xv = v;
        uint32_t qd[6], vd[2];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int d = static_cast<int>((xq + 128) & 255) - 128;
            qd[k] = static_cast<uint32_t>(d) & 255u;
            xq = (xq - d) >> 8;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int d = static_cast<int>((xv + 128) & 255) - 128;
            vd[k] = static_cast<uint32_t>(d) & 255u;
            xv = (xv - d) >> 8;
        }
        const uint32_t d0 = qd[0] | (qd[4] << 8) | (vd[0] << 16), d1 = qd[1] | (qd[5] << 8) | (vd[1] << 16);
        yq[(K + 3) / 4 * 64 + lane] = make_uint4(qd[3], qd[2], d1, d0);
    }
    if (tile + tstep < ntiles) load_y(tile + tstep);

    // 2. limb partials: tiles st = wave + 4 i (i < 4) for all four column tiles,
    // and the 17th limb tile (st = 16) for column tile t = wave. Tiles past
    // ceil((4 L + 4) / 16) read zero table rows, so their partials are 0.
    const int g = lane >> 4, cl = lane & 15;
    long long p0[4][kTiles], p1;
    {
        v4i acc[4][kTiles], acc1 = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int t = 0; t < kTiles; ++t) acc[i][t] = v4i{0, 0, 0, 0};
        const uint32_t* ta = tab + g * kTabRows + 16 * wave + cl;   // slot i at + 64 i, word k at + k
        const uint32_t* ta4 = tab + g * kTabRows + 256 + cl;
        const uint4* yb = yq + g * 64 + cl;                          // column tile t at + 16 t
        const uint4* yx = yq + g * 64 + 16 * wave + cl;
        // unrolled: the accumulators keep fixed registers, the reads immediate offsets
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            // the K-step's table rows through one opaque LDS address each, so
            // that every pair read takes its word offsets as immediates (the
            // compiler otherwise forms one address per pair)
            lds_u32 tk = lds_opaque(ta + 4 * ks * kTabRows);
            lds_u32 tk4 = lds_opaque(ta4 + 4 * ks * kTabRows);
            const uint4* yk = yb + 4 * ks * 64;
            v4i af[4], a4, bf[kTiles], bx;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int k = 0; k < 4; ++k) af[i][k] = static_cast<int>(tk[64 * i + k]);
#pragma unroll
            for (int k = 0; k < 4; ++k) a4[k] = static_cast<int>(tk4[k]);
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                const uint4 u = yk[16 * t];
                bf[t] = v4i{static_cast<int>(u.x), static_cast<int>(u.y), static_cast<int>(u.z), static_cast<int>(u.w)};
            }
            {
                const uint4 u = yx[4 * ks * 64];
                bx = v4i{static_cast<int>(u.x), static_cast<int>(u.y), static_cast<int>(u.z), static_cast<int>(u.w)};
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < kTiles; ++t) acc[i][t] = mfma16(af[i], bf[t], acc[i][t]);
            acc1 = mfma16(a4, bx, acc1);
        }
        int k8 = 1 << 8, k16 = 1 << 16, k24 = 1 << 24;
        asm volatile("" : "+s"(k8), "+s"(k16), "+s"(k24));
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int t = 0; t < kTiles; ++t) p0[i][t] = fold_limb(acc[i][t], k8, k16, k24);
        p1 = fold_limb(acc1, k8, k16, k24);
    }
    __syncthreads();   // y digits and the digit table are dead from here
    // lane (g, cl), tile st, column tile t: limb 4 st + g of coefficient 16 t + cl
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int st = wave + 4 * i;
#pragma unroll
        for (int t = 0; t < kTiles; ++t) part[(4 * st + g) * 64 + 16 * t + cl] = p0[i][t];
    }
    part[(64 + g) * 64 + 16 * wave + cl] = p1;
    __syncthreads();

    // 3. one segment of 17 limbs per wave, one coefficient per lane; limbs past
    // L + 2 are zero and only carry R's sign along. The partials already hold
    // v (-P mod q) - Q q (the quotient group); the offset correction remains.
    const int l0 = kSegMax * wave;
    // limb l of q and of the offset correction at l + 2 (l >= -2), zero
    // outside; loaded per tile (scalar loads), not hoisted out of the tile loop
    const kconst<uint32_t> lc = laundered(P.crt_lc);
    const kconst<uint32_t> lq = lc + 2;
    const kconst<uint32_t> lcorr = lc + 144 + 2;
    auto qlimb = [&](int l) -> uint32_t { return lq[l]; };
    uint32_t x[kSegMax];
    // s's high word is the signed carry into the next limb; it enters the next
    // sum sign-extended, by a multiply-add with an opaque 1
    long long s = 0;
    int one = 1;
    asm volatile("" : "+s"(one));
#pragma unroll
    for (int u = 0; u < kSegMax; ++u) {
        const int l = l0 + u;
        s = part[l * 64 + lane] + static_cast<long long>(lcorr[l]) + static_cast<long long>(static_cast<int>(s >> 32)) * one;
        x[u] = static_cast<uint32_t>(s);
    }
    const long long carry = s >> 32;
    {
        bool uz = true, uo = true;
#pragma unroll
        for (int u = 2; u < kSegMax; ++u) {
            uz = uz && x[u] == 0u;
            uo = uo && x[u] == 0xffffffffu;
        }
        ex_c[wave * 64 + lane] = static_cast<int>(carry);   // |carry| < 2^17
        ex_s[wave * 64 + lane] = lo_summary(static_cast<uint64_t>(x[0]) | (static_cast<uint64_t>(x[1]) << 32));
        ex_f[wave * 64 + lane] = (uz ? 1u : 0u) | (uo ? 2u : 0u);
    }
    __syncthreads();
    int Cin = 0, Ctop = 0;
    {
        int C = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            Cin = w == wave ? C : Cin;
            const uint32_t fl = ex_f[w * 64 + lane];
            C = seg_carry(C, ex_c[w * 64 + lane], ex_s[w * 64 + lane], fl & 1u, (fl & 2u) != 0u);
        }
        Ctop = C;   // R's sign: |R| < 2^(32 L + 1)
    }
    {
        // the segment plus Cin sign-extended, modulo 2^(32 kSegMax): one
        // add-with-carry per limb
        const uint32_t ext = Cin < 0 ? 0xffffffffu : 0u;
        unsigned int c = 0;
        x[0] = __builtin_addc(x[0], static_cast<uint32_t>(Cin), 0u, &c);
#pragma unroll
        for (int u = 1; u < kSegMax; ++u) x[u] = __builtin_addc(x[u], ext, c, &c);
    }
    // the segment against q's: the borrow out of the segment minus q's limbs
    // (one subtract-with-borrow per limb) says below, the OR of the differences
    // equal; +1, 0 or -1
    {
        unsigned int b = 0, diff = 0;
#pragma unroll
        for (int u = 0; u < kSegMax; ++u) diff |= __builtin_subc(x[u], qlimb(l0 + u), b, &b);
        cmpw[wave * 64 + lane] = diff == 0u ? 0 : (b != 0u ? -1 : 1);
    }
    __syncthreads();
    long long sign;
    {
        int cmp = 0;
#pragma unroll
        for (int w = 3; w >= 0; --w) cmp = cmp != 0 ? cmp : cmpw[w * 64 + lane];
        sign = Ctop < 0 ? 1 : (cmp >= 0 ? -1 : 0);   // add q, subtract q, or keep
    }
    // The correction changes a coefficient only when its quotient estimate was
    // one off (rare); with every sign zero it would add 0 q and carry nothing,
    // so a workgroup whose lanes all keep R skips it. The vote: one word per
    // wave in ex_s, whose summaries every wave read before the last barrier;
    // the ballot runs with the whole wave active, outside the lane-0 branch.
    const bool wave_needs = __builtin_amdgcn_ballot_w64(sign != 0) != 0;
    if (lane == 0) ex_s[wave] = wave_needs ? 1 : 0;
    __syncthreads();
    if ((ex_s[0] | ex_s[1] | ex_s[2] | ex_s[3]) != 0) {
    {
        long long cc = 0;
        bool az = true, ao = true;
#pragma unroll
        for (int u = 0; u < kSegMax; ++u) {
            const long long t = static_cast<long long>(x[u]) + sign * static_cast<long long>(qlimb(l0 + u)) + cc;
            x[u] = static_cast<uint32_t>(t);
            cc = t >> 32;
            az = az && x[u] == 0u;
            ao = ao && x[u] == 0xffffffffu;
        }
        // exports of the first resolution are dead (every wave read them before the last barrier)
        ex_c[wave * 64 + lane] = static_cast<int>(cc);
        ex_f[wave * 64 + lane] = (az ? 1u : 0u) | (ao ? 2u : 0u);
    }
    __syncthreads();
    {
        long long C = 0;
#pragma unroll
        for (int w = 0; w < 3; ++w) {
            const uint32_t fl = ex_f[w * 64 + lane];
            const long long over = (C == 1 && (fl & 2u)) ? 1 : ((C == -1 && (fl & 1u)) ? -1 : 0);
            C = w < wave ? ex_c[w * 64 + lane] + over : C;
        }
#pragma unroll
        for (int u = 0; u < kSegMax; ++u) {
            const long long t = static_cast<long long>(x[u]) + C;
            x[u] = static_cast<uint32_t>(t);
            C = t >> 32;
        }
    }
    }
    // 4. the limbs to the stage (the partials are dead), then whole-line stores
    // of the L limbs
#pragma unroll
    for (int u = 0; u < kSegMax; ++u) stage[lane * kStage + l0 + u] = x[u];
    __syncthreads();
    // one coefficient's L words per wave and step (lane = limb)
    uint32_t* dst = c + static_cast<size_t>(nb) * L;
    if (lane < L) {
#pragma unroll 4
        for (uint32_t r = wave; r < nk; r += kWaves) dst[r * L + lane] = stage[r * kStage + lane];
    }
    __syncthreads();   // the stores read the stage (region 0) before the next tile's digits overwrite it
    }
}

template <int KS>
__global__ void __launch_bounds__(64 * kWaves)
k_crt_mfma(const uint32_t* __restrict__ Y, uint32_t* __restrict__ c, DevPlan P, uint32_t k0, uint32_t count) {
    crt_tiles<KS>(Y, c, P, k0, count, blockIdx.x, gridDim.x);
}

// A batch: blockIdx.z = product, with that product's tables; every product's
// K needs KS K-steps.
template <int KS>
__global__ void __launch_bounds__(64 * kWaves) k_crt_mfma_batch(DevPlan P, BatchArgs ba) {
    const int i = blockIdx.z;
    const auto d = as_const(ba.d[i]);
    DevPlan Q = P;
    Q.K = d->K;
    Q.L = d->L;
    Q.crt_tab = d->crt_tab;
    Q.crt_fc = d->crt_fc;
    Q.nq = d->nq;
    Q.crt_lc = d->crt_lc;
    crt_tiles<KS>(ba.A[i], ba.c[i], Q, 0, P.N, blockIdx.x, gridDim.x);
}

// Balanced base-256 digits of a signed value |x| < 2^31.
void balanced4(long long x, int8_t out[4]) {
    for (int i = 0; i < 4; ++i) {
        const long long r = ((x + 128) % 256 + 256) % 256 - 128;
        out[i] = static_cast<int8_t>(r);
        x = (x - r) / 256;
    }
    if (x != 0) throw std::logic_error("balanced4: value out of range");
}

}  // namespace

void prepare_mfma(const HostPlan& h, DevPlan& d, const std::function<void*(const void*, size_t)>& upload) {
    const int K = h.K, L = h.L;
    if (L > 4 * kDecMaxKSteps * 4 || ((K + 3) / 4 + 1 + 3) / 4 > kCrtMaxKSteps) {
        throw std::runtime_error("point too large for the matrix-core tables");
    }
    // Decompose tables: T_{j,t} = f_j 256^t 2^32 mod p_j with f_j = 1 (the plain
    // tables) or f_j = scale_j (the b side: b scale_j mod p_j); the factor 2^32
    // is what the kernel's Montgomery step takes away.
    const int KT = (4 * L + 63) / 64;
    const int KG = (K + 3) / 4;
    auto dec_tables = [&](bool scaled, const int8_t** digits, const int2** init) {
        std::vector<int8_t> dec(static_cast<size_t>(KG) * KT * 64 * 16, 0);
        std::vector<int2> dinit(K);
        for (int j = 0; j < K; ++j) {
            const uint64_t p = h.p[j];
            uint64_t T = (scaled ? h.scale[j] % p : 1) * ((uint64_t(1) << 32) % p) % p, sum = 0;
            for (int t = 0; t < 4 * L; ++t) {
                sum = (sum + T) % p;
                const long long centred =
                    T <= p / 2 ? static_cast<long long>(T) : static_cast<long long>(T) - static_cast<long long>(p);
                int8_t e[4];
                balanced4(centred, e);
                const int ks = t / 64, g = (t % 64) / 16, byte = t % 16;
                for (int b = 0; b < 4; ++b) {
                    const int row = 4 * (j % 4) + b;
                    const int lane = 16 * g + row;
                    dec[((static_cast<size_t>(j / 4) * KT + ks) * 64 + lane) * 16 + byte] = e[b];
                }
                T = T * 256 % p;
            }
            // The bytes enter as u - 128, so the sums miss 128 sum_t T; with the
            // offset 2^16 c (c = 2^30 + 2^24) that keeps the folded sum positive,
            // the accumulators must add d = 128 sum_t T - 2^16 c (mod p), taken
            // centred (|d| < p/2 < 2^30) and split as d = d_lo + 2^16 d_hi with
            // d_lo in [0, 2^16): digit 0 starts at d_lo, digit 2 at d_hi + c.
            const long long c = (1ll << 30) + (1ll << 24);
            long long dd = static_cast<long long>((sum * 128 % p + p - (static_cast<uint64_t>(c) << 16) % p) % p);
            if (dd > static_cast<long long>(p / 2)) dd -= static_cast<long long>(p);
            const long long d_hi = dd >= 0 ? dd >> 16 : -((-dd + 65535) >> 16);
            const long long d_lo = dd - (d_hi << 16);
            if (d_lo < 0 || d_lo >= 65536) throw std::logic_error("decompose offsets");
            dinit[j] = make_int2(static_cast<int>(d_lo), static_cast<int>(d_hi + c));
        }
        *digits = static_cast<const int8_t*>(upload(dec.data(), dec.size()));
        *init = static_cast<const int2*>(upload(dinit.data(), dinit.size() * sizeof(int2)));
    };
    dec_tables(false, &d.dec_digits, &d.dec_init);
    dec_tables(true, &d.dec_digits_b, &d.dec_init_b);
    d.dec_ksteps = KT;
    // i / D = mulhi(i, floor(2^32 / D) + 1) for i < 2^12 and 2 <= D <= 64
    // (every such i and D checked exhaustively): D = L / 4 where the kernel divides
    // 16-byte unit indices (L a multiple of 4, at least 8), else D = L (word
    // indices; L = 1 divides by nothing)
    const uint64_t D = (L % 4 == 0 && L >= 8) ? static_cast<uint64_t>(L / 4) : static_cast<uint64_t>(std::max(L, 2));
    d.dec_magic = static_cast<uint32_t>((uint64_t(1) << 32) / D + 1);

    // CRT tables: the prime groups, then the quotient group (column KJ).
    const int KS = ((K + 3) / 4 + 1 + 3) / 4;
    const int ST = (4 * L + 4 + 15) / 16;
    const int nd = 4 * L + 1;
    std::vector<int8_t> m(static_cast<size_t>(K) * nd);
    for (int j = 0; j < K; ++j) balanced_digits(h.M.data() + static_cast<size_t>(j) * L, L, m.data() + static_cast<size_t>(j) * nd, nd);
    if (ST > 17) throw std::logic_error("CRT limb tiles exceed the kernel's 17");
    const int KJ = (K + 3) / 4;
    if (crt_tab_words(KS) > 4u * kTabLoads * 64 * kWaves) throw std::logic_error("CRT digit table too large");
    std::vector<uint32_t> tab(crt_tab_words(KS), 0);
    for (int J = 0; J < KJ; ++J)
        for (int R = 3; R < kTabRows && R - 3 < nd; ++R)
            for (int r = 0; r < 4 && 4 * J + r < K; ++r)
                tab[static_cast<size_t>(J) * kTabRows + R] |=
                    static_cast<uint32_t>(static_cast<uint8_t>(m[static_cast<size_t>(4 * J + r) * nd + (R - 3)]))
                    << (8 * r);
    {
        // -q as balanced digits: those of 2^(32 (L + 1)) - q over 4 (L + 1) + 1
        // digits, less 256^(4 (L + 1)) at the top (a top digit of 0 or 1 becomes -1 or 0)
        std::vector<uint32_t> tc(L + 1, 0);
        uint64_t borrow = 1;   // ~q + 1 over L + 1 limbs
        for (int l = 0; l <= L; ++l) {
            const uint64_t s = static_cast<uint64_t>(~(l < L ? h.q[l] : 0u)) + borrow;
            tc[l] = static_cast<uint32_t>(s);
            borrow = s >> 32;
        }
        const int nq = 4 * (L + 1) + 1;
        std::vector<int8_t> negq(nq), pn(nd);
        balanced_digits(tc.data(), L + 1, negq.data(), nq);
        negq[nq - 1] = static_cast<int8_t>(negq[nq - 1] - 1);
        balanced_digits(h.negvP.data() + L, L, pn.data(), nd);   // row v = 1: -P mod q
        auto put = [&](int digit, int r, int8_t v) {
            if (digit + 3 >= kTabRows) throw std::logic_error("CRT quotient digits exceed the table");
            tab[static_cast<size_t>(KJ) * kTabRows + digit + 3] |= static_cast<uint32_t>(static_cast<uint8_t>(v)) << (8 * r);
        };
        for (int s = 0; s < nq; ++s) {
            put(s, 0, negq[s]);       // Q's digits 0-3 times -q
            put(s + 4, 1, negq[s]);   // digits 4-5 times -q 2^32
        }
        for (int s = 0; s < nd; ++s) put(s, 2, pn[s]);   // v times -P mod q
    }
    // 0x80808080 * sum_j M_j, exactly (L + 2 limbs).
    std::vector<uint64_t> acc(L + 2, 0);
    for (int j = 0; j < K; ++j)
        for (int l = 0; l < L; ++l) acc[l] += h.M[static_cast<size_t>(j) * L + l];
    std::vector<uint32_t> sumM(L + 2, 0);
    uint64_t carry = 0;
    for (int l = 0; l < L + 2; ++l) {
        const uint64_t s = acc[l] + carry;
        sumM[l] = static_cast<uint32_t>(s);
        carry = s >> 32;
    }
    std::vector<uint32_t> ccorr(L + 2, 0);
    carry = 0;
    for (int l = 0; l < L + 2; ++l) {
        const uint64_t s = static_cast<uint64_t>(sumM[l]) * 0x80808080u + carry;
        ccorr[l] = static_cast<uint32_t>(s);
        carry = s >> 32;
    }
    if (carry) throw std::logic_error("CRT correction overflows L + 2 limbs");
    d.crt_tab = static_cast<const uint32_t*>(upload(tab.data(), tab.size() * sizeof(uint32_t)));
    constexpr int kFc = 4 * 4 * kCrtMaxKSteps;
    std::vector<double> fc(2 * kFc, 0.0);
    for (int j = 0; j < K; ++j) {
        fc[j] = h.p_recip[j];
        fc[kFc + j] = h.mq[j];
    }
    d.crt_fc = static_cast<const double*>(upload(fc.data(), fc.size() * sizeof(double)));
    std::vector<uint32_t> lc(3 * 72, 0);
    for (int l = 0; l < L; ++l) {
        lc[l + 2] = h.q[l];
        lc[72 + l + 2] = h.negvP[static_cast<size_t>(L) + l];   // row v = 1: -P mod q
    }
    for (int l = 0; l < L + 2; ++l) lc[144 + l + 2] = ccorr[l];
    d.crt_lc = static_cast<const uint32_t*>(upload(lc.data(), lc.size() * sizeof(uint32_t)));
    d.crt_corr = static_cast<const uint32_t*>(upload(ccorr.data(), ccorr.size() * sizeof(uint32_t)));
    d.crt_ksteps = KS;
    d.crt_stiles = ST;
}

void launch_decompose_mfma(const DevPlan& P, const uint32_t* limbs, uint32_t* res, hipStream_t st, bool lazy_out) {
    launch_decompose_mfma_range(P, limbs, res, 0, P.N, st, lazy_out);
}

namespace {

template <int KT>
void dec_launch(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0, uint32_t count, hipStream_t st,
                bool lazy_out, const BatchArgs* ba, int nprod) {
    const size_t lds = sizeof(uint32_t) * 16 * kTiles * (16 * KT + 4) + 16 * static_cast<size_t>(P.K);
    const dim3 block(64 * kWaves);
    if (ba) {
        const dim3 grid((P.N + 16 * kTiles - 1) / (16 * kTiles), 1, 2 * nprod);
        if (lazy_out)
            launch(k_decompose_mfma_batch<KT, true>, grid, block, lds, st, P, *ba);
        else
            launch(k_decompose_mfma_batch<KT, false>, grid, block, lds, st, P, *ba);
    } else {
        const dim3 grid((count + 16 * kTiles - 1) / (16 * kTiles));
        if (lazy_out)
            launch(k_decompose_mfma<KT, true>, grid, block, lds, st, limbs, res, P, k0, count);
        else
            launch(k_decompose_mfma<KT, false>, grid, block, lds, st, limbs, res, P, k0, count);
    }
}

void dec_dispatch(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0, uint32_t count, hipStream_t st,
                  bool lazy_out, const BatchArgs* ba = nullptr, int nprod = 0) {
    if (!P.dec_digits) throw std::runtime_error("matrix-core tables not prepared");
    switch (P.dec_ksteps) {
        case 1: dec_launch<1>(P, limbs, res, k0, count, st, lazy_out, ba, nprod); return;
        case 2: dec_launch<2>(P, limbs, res, k0, count, st, lazy_out, ba, nprod); return;
        case 3: dec_launch<3>(P, limbs, res, k0, count, st, lazy_out, ba, nprod); return;
        case 4: dec_launch<4>(P, limbs, res, k0, count, st, lazy_out, ba, nprod); return;
        default: throw std::runtime_error("L out of range");
    }
}

}  // namespace

void launch_decompose_mfma_range(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0,
                                 uint32_t count, hipStream_t st, bool lazy_out) {
    if (count == 0) return;
    dec_dispatch(P, limbs, res, k0, count, st, lazy_out && P.lazy != 0);
}

void launch_crt_mfma(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st) {
    launch_crt_mfma_range(P, Y, c, 0, P.N, st);
}

// Workgroups of the CRT kernel that fit on the current device at once: its LDS
// allows floor(160 KiB / lds) per CU (the workgroups are persistent and loop
// over 64-coefficient tiles).
uint32_t crt_slots(size_t lds) {
    static thread_local int dev = -1, cus = 0;
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    if (d != dev) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || cus <= 0)
            throw std::runtime_error("hipDeviceGetAttribute failed");
        dev = d;
    }
    const size_t per_cu = std::max<size_t>(1, (160u * 1024u) / lds);
    return static_cast<uint32_t>(per_cu * static_cast<size_t>(cus));
}

namespace {

// The residues are read through a buffer whose byte range must fit in 31 bits.
void check_crt_range(const DevPlan& P) {
    if (!P.crt_tab) throw std::runtime_error("matrix-core tables not prepared");
    if (4ull * static_cast<uint64_t>(P.K) * P.N >= (1ull << 31)) throw std::runtime_error("K x N too large for the CRT");
    if (P.crt_ksteps < 1 || P.crt_ksteps > kCrtMaxKSteps) throw std::runtime_error("K out of range for the CRT");
}

template <int KS>
void crt_launch(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count, hipStream_t st,
                const BatchArgs* ba, int nprod) {
    const size_t lds = sizeof(uint32_t) * crt_lds_words(KS);
    if (ba) {
        const uint32_t per_product = std::max<uint32_t>(1, crt_slots(lds) / static_cast<uint32_t>(nprod));
        const dim3 grid(std::min((P.N + 63) / 64, per_product), 1, nprod);
        launch(k_crt_mfma_batch<KS>, grid, dim3(64 * kWaves), lds, st, P, *ba);
    } else {
        const dim3 grid(std::min((count + 63) / 64, crt_slots(lds)));
        launch(k_crt_mfma<KS>, grid, dim3(64 * kWaves), lds, st, Y, c, P, k0, count);
    }
}

void crt_dispatch(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count, hipStream_t st,
                  const BatchArgs* ba = nullptr, int nprod = 0) {
    check_crt_range(P);
    switch (P.crt_ksteps) {
        case 1: crt_launch<1>(P, Y, c, k0, count, st, ba, nprod); return;
        case 2: crt_launch<2>(P, Y, c, k0, count, st, ba, nprod); return;
        case 3: crt_launch<3>(P, Y, c, k0, count, st, ba, nprod); return;
        case 4: crt_launch<4>(P, Y, c, k0, count, st, ba, nprod); return;
        case 5: crt_launch<5>(P, Y, c, k0, count, st, ba, nprod); return;
        case 6: crt_launch<6>(P, Y, c, k0, count, st, ba, nprod); return;
        case 7: crt_launch<7>(P, Y, c, k0, count, st, ba, nprod); return;
        case 8: crt_launch<8>(P, Y, c, k0, count, st, ba, nprod); return;
        case 9: crt_launch<9>(P, Y, c, k0, count, st, ba, nprod); return;
        case 10: crt_launch<10>(P, Y, c, k0, count, st, ba, nprod); return;
        default: crt_launch<11>(P, Y, c, k0, count, st, ba, nprod); return;
    }
}

}  // namespace

void launch_crt_mfma_range(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count,
                           hipStream_t st) {
    if (count == 0) return;
    crt_dispatch(P, Y, c, k0, count, st);
}

// Always a product's decomposition: the forward NTT follows, so the residues
// may stay in [0, 2p) when the plan's butterflies accept that.
void launch_decompose_mfma_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st) {
    dec_dispatch(P, nullptr, nullptr, 0, P.N, st, P.lazy != 0, &ba, count);
}

namespace {

template <int R1, int KT>
void fc_launch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st, int side) {
    // every product of the launch takes its own K; the grid covers the longest
    const int groups = (P.K + 3) / 4;
    const int gps = 4 * ((groups + 3) / 4);
    const dim3 grid((P.N / R1) / kFcCols, (groups + gps - 1) / gps, side < 0 ? 2 * count : count);
    const size_t lds = fc_lds_bytes<R1>(P.K);
    if (P.lazy)
        launch(k_decompose_cols_batch<R1, KT, 1>, grid, dim3(32 * R1), lds, st, P, ba, side, gps);
    else
        launch(k_decompose_cols_batch<R1, KT, 0>, grid, dim3(32 * R1), lds, st, P, ba, side, gps);
}

template <int R1>
void fc_dispatch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st, int side) {
    switch (P.dec_ksteps) {
        case 1: fc_launch<R1, 1>(P, ba, count, st, side); return;
        case 2: fc_launch<R1, 2>(P, ba, count, st, side); return;
        case 3: fc_launch<R1, 3>(P, ba, count, st, side); return;
        case 4: fc_launch<R1, 4>(P, ba, count, st, side); return;
        default: throw std::runtime_error("L out of range");
    }
}

}  // namespace

void launch_decompose_cols_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st, int side) {
    if (!fused_cols_ok(P)) throw std::logic_error("fused column stages need N = 32768 or 65536 and the matrix-core tables");
    if (P.logR == 8)
        fc_dispatch<16>(P, ba, count, st, side);
    else
        fc_dispatch<8>(P, ba, count, st, side);
}

void launch_crt_mfma_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st) {
    crt_dispatch(P, nullptr, nullptr, 0, P.N, st, &ba, count);
}

}  // namespace pm
