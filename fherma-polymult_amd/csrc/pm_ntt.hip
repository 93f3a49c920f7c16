// Register-resident NTT passes for sub-transform lengths M = 2^m, 4 <= m <= 8.
//
// Same transform as the generic kernels in pm_device.hip (the merged
// Cooley-Tukey NTT, twiddle index (N + k) / (2t) for a butterfly of distance t
// whose lower element is at k), scheduled for wave64: every thread owns 16
// elements of one sub-transform (a column of R elements with stride S, or a row
// of S contiguous ones) and runs several stages on them in registers, with a
// single LDS exchange per pass:
//   phase A  stages of distance 2^(m-1) .. 16: register i GA + k holds element
//            u = t + (M/16) i + 16 k of the sub-transform (GA = M/16 per group)
//   phase B  stages of distance 8 .. 1: register k holds element u = 16 t + k
// (forward runs A then B, inverse B then A). In both phases a butterfly block's
// twiddle index is (B >> (e + 1)) plus a block number known at compile time
// (phase A) or from t (phase B), where 2^e is the distance and B = R for the
// column pass, B = N + rho S for row rho.
//
// LDS holds element u of a sub-transform at u + u/16 (one pad word per 16), so
// that the phase B reads (16 lanes with stride 16) hit distinct banks.
#include "pm_device.h"

#include "pm_arith.h"
#include "pm_bfly.h"

#include <stdexcept>

namespace pm {

void launch_ntt_cols_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
void launch_ntt_rows_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
void launch_ntt_rows_fused_generic(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st);

namespace {

// Twiddle sources: tw(e, b) is the twiddle of block b (counted from the start
// of the sub-transform) of the stage with distance 2^e.
// tw.b(e, t, blk): phase B's twiddle for thread t, block blk of its 16 elements.
// The full K x N tables hold Shoup pairs (forward) or Montgomery pairs (inverse).
template <bool kMontTw>
struct TableTw {   // the full K x N table: index (B >> (e + 1)) + b
    static constexpr bool kMont = kMontTw;
    const uint2* __restrict__ tab;
    uint32_t B;
    __device__ __forceinline__ uint2 operator()(int e, uint32_t b) const { return tab[(B >> (e + 1)) + b]; }
    __device__ __forceinline__ uint2 b(int e, uint32_t t, int blk) const {
        return (*this)(e, (t << (3 - e)) + blk);
    }
};

// Row rho's twiddle of stage e, block b, as a Montgomery pair: Z[rho][e] is
// stored as Z 2^32 mod p with its Shoup companion, so its Shoup product with
// Om[b] is w' = Z Om 2^32 mod p, canonical, and w' pinv (pinv = p^-1 mod 2^32)
// completes the pair.
struct GenTw {
    static constexpr bool kMont = true;
    const uint2* __restrict__ z;
    const uint32_t* __restrict__ om;
    uint32_t p, pinv;
    __device__ __forceinline__ uint2 operator()(int e, uint32_t b) const {
        const uint32_t w = mul_shoup(om[b], z[e].x, z[e].y, p);
        return make_uint2(w, w * pinv);
    }
    __device__ __forceinline__ uint2 b(int e, uint32_t t, int blk) const {
        return (*this)(e, (t << (3 - e)) + blk);
    }
};

// Phase A twiddles of one row, precomputed in LDS: stage e, block b at
// 2^(m-1-e) - 1 + b (2^(m-4) - 1 entries).
template <int LOGM, bool kMontTw>
struct LdsTwA {
    static constexpr bool kMont = kMontTw;
    const uint2* tab;
    __device__ __forceinline__ uint2 operator()(int e, uint32_t b) const {
        return tab[(1u << (LOGM - 1 - e)) - 1 + b];
    }
};

template <int LOGM, class TW>
__device__ __forceinline__ void fill_row_phase_a(uint2* tab, const TW& g, uint32_t t) {
    constexpr uint32_t n = (1u << (LOGM - 4)) - 1;
    if (t < n) {
        const uint32_t m = 31 - __clz(t + 1);   // floor(log2(t + 1))
        tab[t] = g(LOGM - 1 - m, t + 1 - (1u << m));
    }
}

// The row pass's twiddles for row rho of prime j: generated from the per-row
// and per-prime factors (kGen), or read from the full table, whose index for
// row rho is ((N + rho S) >> (e + 1)) + b.
template <int LOGM, bool kGen, bool fwd>
__device__ __forceinline__ auto row_tw(const DevPlan& P, int j, uint32_t rho, const PrimeConst& pc) {
    if constexpr (kGen) {
        const uint32_t R = 1u << P.logR;
        return GenTw{(fwd ? P.rz : P.rzi) + (static_cast<size_t>(j) * R + rho) * LOGM,
                     (fwd ? P.rom : P.romi) + static_cast<size_t>(j) * ((1u << LOGM) / 2), pc.p, 0u - pc.pneg};
    } else {
        return TableTw<!fwd>{(fwd ? P.tw : P.itw) + static_cast<size_t>(j) * P.N, P.N + (rho << LOGM)};
    }
}

// Phase A on one or two register sets (kTwo: x and y share the twiddles).
template <int LOGM, bool kFwd, bool kTwo, int kMode, class TW>
__device__ __forceinline__ void phase_a(uint32_t (&x)[16], uint32_t (&y)[16], const TW& tw, uint32_t p) {
    constexpr int GA = 1 << (LOGM - 4), NG = 16 / GA;
#pragma unroll
    for (int s = 0; s < LOGM - 4; ++s) {
        const int e = kFwd ? LOGM - 1 - s : 4 + s;
        const int d = 1 << (e - 4);
#pragma unroll
        for (int blk = 0; blk < GA / (2 * d); ++blk) {
            const uint2 w = tw(e, blk);
#pragma unroll
            for (int i = 0; i < NG; ++i) {
#pragma unroll
                for (int k = blk * 2 * d; k < blk * 2 * d + d; ++k) {
                    bfly<kFwd, kMode, TW::kMont>(x[i * GA + k], x[i * GA + k + d], w, p);
                    if (kTwo) bfly<kFwd, kMode, TW::kMont>(y[i * GA + k], y[i * GA + k + d], w, p);
                }
            }
        }
    }
}

// The column pass's twiddles (the first R pairs of the prime's table, all
// that its stages use) through a buffer: phase A's are
// wave-uniform (scalar loads), phase B's index (B >> (e + 1)) + (t << (3 - e)) + blk
// splits into a lane part (t << (3 - e), four values fixed per thread) and a
// uniform part.
template <bool kMontTw>
struct BufTw {
    static constexpr bool kMont = kMontTw;
    const uint2* __restrict__ tab;
    Buf buf;
    uint32_t B, vt[4];
    __device__ __forceinline__ BufTw(const uint2* tab_, uint32_t entries, uint32_t B_, uint32_t t)
        : tab(tab_), buf(tab_, 8u * entries), B(B_) {
#pragma unroll
        for (int e = 0; e < 4; ++e) vt[e] = 8u * (t << (3 - e));
    }
    __device__ __forceinline__ uint2 operator()(int e, uint32_t b) const { return tab[(B >> (e + 1)) + b]; }
    __device__ __forceinline__ uint2 b(int e, uint32_t, int blk) const {
        return buf.ld2(vt[e], 8u * ((B >> (e + 1)) + blk));
    }
};

// Column pass: C = 2^LOGC adjacent columns of one prime, R = M rows each. LDS
// slot padded(u) of column c is at padded(u) C + c; with t < TPS and
// TPS NG = 16, padded(t + TPS i + 16 k) = t + TPS i + 17 k and
// padded(16 t + k) = 17 t + k, so every slot is a lane base plus a constant.
// kMode 1: the forward pass leaves its output in [0, 2p).
template <int LOGM, bool kFwd, int LOGC, int kMode>
__device__ __forceinline__ void col_pass(uint32_t* __restrict__ X, const DevPlan& P) {
    extern __shared__ uint32_t lds[];
    constexpr uint32_t M = 1u << LOGM, TPS = M / 16, GA = TPS, NG = 16 / GA, C = 1u << LOGC;
    const uint32_t c = threadIdx.x & (C - 1), t = threadIdx.x >> LOGC;
    const int j = blockIdx.y;
    const uint32_t S = 1u << P.logS;
    // element (row u, this column) at byte 4 (u S + column): lane part + uniform part
    const Buf x(X + static_cast<size_t>(j) * P.N, 4u * P.N);
    const uint32_t col = (blockIdx.x << LOGC) + c;
    const uint32_t va = 4u * (t * S + col), vb = 4u * (16u * t * S + col);   // phase A / phase B rows
    const BufTw<!kFwd> tw((kFwd ? P.ctw : P.citw) + static_cast<size_t>(j) * M, M, M, t);
    uint32_t* la = lds + t * C + c;          // phase A slots: + (TPS i + 17 k) C
    uint32_t* lb = lds + 17u * t * C + c;    // phase B slots: + k C
    const uint32_t p = P.pc[j].p;
    uint32_t r[16];
    if (kFwd) {
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) r[i * GA + k] = x.ld(va, 4u * (TPS * i + 16 * k) * S);
        phase_a<LOGM, true, false, kMode>(r, r, tw, p);
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) la[(TPS * i + 17 * k) * C] = r[i * GA + k];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) r[k] = lb[k * C];
        phase_b<true, false, kMode>(r, r, tw, t, p);
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) x.st(r[k], vb, 4u * k * S);
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) r[k] = x.ld(vb, 4u * k * S);
        phase_b<false, false, kMode>(r, r, tw, t, p);
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) lb[k * C] = r[k];
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) r[i * GA + k] = la[(TPS * i + 17 * k) * C];
        phase_a<LOGM, false, false, kMode>(r, r, tw, p);
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) x.st(r[i * GA + k], va, 4u * (TPS * i + 16 * k) * S);
    }
}

template <int LOGM, bool kFwd, int LOGC, int kMode>
__global__ void __launch_bounds__(512) k_ntt_col_reg(uint32_t* __restrict__ X, DevPlan P) {
    col_pass<LOGM, kFwd, LOGC, kMode>(X, P);
}

// A batch: forward, blockIdx.z = 2 (product) + (0: A, 1: B); inverse, blockIdx.z = product (A).
template <int LOGM, bool kFwd, int LOGC, int kMode>
__global__ void __launch_bounds__(512) k_ntt_col_batch(DevPlan P, BatchArgs ba) {
    const int i = kFwd ? blockIdx.z >> 1 : blockIdx.z;
    if (static_cast<int>(blockIdx.y) >= ba.K[i]) return;   // primes past this product's K
    col_pass<LOGM, kFwd, LOGC, kMode>(kFwd && (blockIdx.z & 1) ? ba.B[i] : ba.A[i], P);
}

// Row pass alone (forward or inverse): G = 2^logG rows of M = S elements.
// Twiddles are generated per row (kGen) or read from the full table; phase
// A's are staged once per row into LDS by the row's first threads. The output
// is canonical; the forward pass accepts input in [0, 2p) (kMode 1).
template <int LOGM, bool kFwd, bool kGen, int kMode>
__global__ void __launch_bounds__(512)
k_ntt_row_reg(uint32_t* __restrict__ X, DevPlan P, uint32_t logG) {
    extern __shared__ uint32_t lds[];
    constexpr uint32_t M = 1u << LOGM, TPS = M / 16, GA = TPS, NG = 16 / GA, ROW = M + M / 16;
    const uint32_t t = threadIdx.x & (TPS - 1), rl = threadIdx.x / TPS;
    const uint32_t G = 1u << logG;
    const uint32_t rho = (blockIdx.x << logG) + rl;
    const int j = blockIdx.y;
    uint32_t* x = X + static_cast<size_t>(j) * P.N + static_cast<size_t>(rho) * M;
    const PrimeConst pc = P.pc[j];
    const auto gen = row_tw<LOGM, kGen, kFwd>(P, j, rho, pc);
    uint32_t* l = lds + rl * ROW;
    uint2* twa = reinterpret_cast<uint2*>(lds + G * ROW) + rl * GA;
    fill_row_phase_a<LOGM>(twa, gen, t);
    const LdsTwA<LOGM, decltype(gen)::kMont> ta{twa};
    uint32_t r[16];
    if (kFwd) {
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) r[i * GA + k] = x[t + TPS * i + 16 * k];
        __syncthreads();
        phase_a<LOGM, true, false, kMode>(r, r, ta, pc.p);
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) l[padded(t + TPS * i + 16 * k)] = r[i * GA + k];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) r[k] = l[padded(16 * t + k)];
        phase_b<true, false, kMode>(r, r, gen, t, pc.p);
        if (kMode == 1) {
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k) r[k] = csub(r[k], pc.p);
        }
        // phase B layout: 16 consecutive words per thread
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) x[16 * t + k] = r[k];
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) r[k] = x[16 * t + k];
        phase_b<false, false, kMode>(r, r, gen, t, pc.p);
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) l[padded(16 * t + k)] = r[k];
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) r[i * GA + k] = l[padded(t + TPS * i + 16 * k)];
        phase_a<LOGM, false, false, kMode>(r, r, ta, pc.p);
#pragma unroll
        for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
            for (uint32_t k = 0; k < GA; ++k) x[t + TPS * i + 16 * k] = r[i * GA + k];
    }
}

// In a product, B was decomposed with the b-side tables (b scale_j, scale_j =
// N^-1 (P/p_j)^-1 2^32), so the Montgomery product a b scale_j 2^-32 is already
// the pointwise term a b N^-1 (P/p_j)^-1.
__device__ __forceinline__ uint32_t pointwise_mul(uint32_t a, uint32_t b, const PrimeConst& pc) {
    return mont_mul(a, b, pc.p, pc.pneg);
}

// Row pass fused: forward rows of A and B (sharing twiddle loads), the pointwise
// product with N^-1 (P/p_j)^-1 into A, inverse rows of A. kMode 1: input in
// [0, 2p); a brought to [0, p) keeps a b below 2 p^2 < 2^32 p, which the
// canonical Montgomery product needs.
// The 16 values per thread of a row of A and of B (phase A layout) at `row`.
template <int LOGM>
__device__ __forceinline__ void load_row_pair(const uint32_t* __restrict__ A, const uint32_t* __restrict__ Bm,
                                              size_t row, uint32_t t, uint32_t (&a)[16], uint32_t (&b)[16]) {
    constexpr uint32_t M = 1u << LOGM, TPS = M / 16, GA = TPS, NG = 16 / GA;
#pragma unroll
    for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
        for (uint32_t k = 0; k < GA; ++k) {
            a[i * GA + k] = A[row + t + TPS * i + 16 * k];
            b[i * GA + k] = Bm[row + t + TPS * i + 16 * k];
        }
}

// One row of a product once its values are in registers: forward rows of A and
// B (sharing twiddle loads), the pointwise product into A, inverse row of A,
// stored at `row`. la: the row's exchange buffer, twa: its phase-A twiddle slots.
template <int LOGM, bool kGen, int kMode>
__device__ __forceinline__ void row_item(uint32_t* __restrict__ A, const DevPlan& P, int j, uint32_t rho, size_t row,
                                         uint32_t t, uint32_t* la, uint2* twa, uint32_t (&a)[16], uint32_t (&b)[16]) {
    constexpr uint32_t M = 1u << LOGM, TPS = M / 16, GA = TPS, NG = 16 / GA;
    const PrimeConst pc = P.pc[j];
    const auto fw = row_tw<LOGM, kGen, true>(P, j, rho, pc);
    const auto iw = row_tw<LOGM, kGen, false>(P, j, rho, pc);
    fill_row_phase_a<LOGM>(twa, fw, t);
    const LdsTwA<LOGM, decltype(fw)::kMont> fa{twa};
    const LdsTwA<LOGM, decltype(iw)::kMont> ia{twa};
    wave_sync();
    phase_a<LOGM, true, true, kMode>(a, b, fa, pc.p);
#pragma unroll
    for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
        for (uint32_t k = 0; k < GA; ++k) la[padded(t + TPS * i + 16 * k)] = a[i * GA + k];
    wave_sync();
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) a[k] = la[padded(16 * t + k)];
    wave_sync();
#pragma unroll
    for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
        for (uint32_t k = 0; k < GA; ++k) la[padded(t + TPS * i + 16 * k)] = b[i * GA + k];
    wave_sync();
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) b[k] = la[padded(16 * t + k)];
    phase_b<true, true, kMode>(a, b, fw, t, pc.p);
    if (kMode == 1) {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) a[k] = pointwise_mul(csub(a[k], pc.p), b[k], pc);
    } else {
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) a[k] = pointwise_mul(a[k], b[k], pc);
    }
    phase_b<false, false, kMode>(a, a, iw, t, pc.p);
    // Each thread rewrites only the LDS words it read above (b's, phase B layout).
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) la[padded(16 * t + k)] = a[k];
    // the forward phase A read its twiddle slots before the syncs above
    fill_row_phase_a<LOGM>(twa, iw, t);
    wave_sync();
#pragma unroll
    for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
        for (uint32_t k = 0; k < GA; ++k) a[i * GA + k] = la[padded(t + TPS * i + 16 * k)];
    phase_a<LOGM, false, false, kMode>(a, a, ia, pc.p);
#pragma unroll
    for (uint32_t i = 0; i < NG; ++i)
#pragma unroll
        for (uint32_t k = 0; k < GA; ++k) A[row + t + TPS * i + 16 * k] = a[i * GA + k];
}

// Row pass fused: forward rows of A and B, the pointwise product with
// N^-1 (P/p_j)^-1 into A, inverse rows of A; G = 2^logG rows per workgroup.
// kMode 1: input in [0, 2p); a brought to [0, p) keeps a b below 2 p^2 < 2^32 p,
// which the canonical Montgomery product needs. One exchange buffer per row,
// used by a and then b; one phase-A twiddle slot set per row, forward first,
// refilled with the inverse ones before the inverse phase A (19 KB of LDS per
// 16 rows).
template <int LOGM, bool kGen, int kMode>
__device__ __forceinline__ void rows_fused(uint32_t* __restrict__ A, const uint32_t* __restrict__ Bm, const DevPlan& P,
                                           uint32_t logG) {
    extern __shared__ uint32_t lds[];
    constexpr uint32_t M = 1u << LOGM, TPS = M / 16, GA = TPS, ROW = M + M / 16;
    const uint32_t t = threadIdx.x & (TPS - 1), rl = threadIdx.x / TPS;
    const uint32_t G = 1u << logG;
    const uint32_t rho = (blockIdx.x << logG) + rl;
    const int j = blockIdx.y;
    const size_t row = static_cast<size_t>(j) * P.N + static_cast<size_t>(rho) * M;
    uint32_t* la = lds + rl * ROW;
    uint2* twa = reinterpret_cast<uint2*>(lds + G * ROW) + rl * GA;
    uint32_t a[16], b[16];
    load_row_pair<LOGM>(A, Bm, row, t, a, b);
    row_item<LOGM, kGen, kMode>(A, P, j, rho, row, t, la, twa, a, b);
}

template <int LOGM, bool kGen, int kMode>
__global__ void __launch_bounds__(512)
k_ntt_rows_fused_reg(uint32_t* __restrict__ A, const uint32_t* __restrict__ Bm, DevPlan P, uint32_t logG) {
    rows_fused<LOGM, kGen, kMode>(A, Bm, P, logG);
}

// A batch: blockIdx.z = product.
template <int LOGM, bool kGen, int kMode>
__global__ void __launch_bounds__(512) k_ntt_rows_fused_batch(DevPlan P, BatchArgs ba, uint32_t logG) {
    if (static_cast<int>(blockIdx.y) >= ba.K[blockIdx.z]) return;   // primes past this product's K
    rows_fused<LOGM, kGen, kMode>(ba.A[blockIdx.z], ba.B[blockIdx.z], P, logG);
}

// The forward column stages of distance S .. 8 S at rows rho0 .. rho0 + 15 (a
// block of 16 S elements, rho0 a multiple of 16): global index
// (N + k) / (2 S 2^e) = (R + rho0 + i) >> (e + 1) for the element k of block row
// i, which is ((R + rho0) >> (e + 1)) + (i >> (e + 1)), phase B's form for
// t = 0 and B = R + rho0. Uniform over the workgroup: scalar loads from the
// first R pairs of the prime's forward table.
struct MidTw {
    static constexpr bool kMont = false;
    kconst_ptr<uint64_t> tab;
    uint32_t B;
    __device__ __forceinline__ uint2 b(int e, uint32_t, int blk) const {
        const uint64_t w = tab[(B >> (e + 1)) + blk];
        return make_uint2(static_cast<uint32_t>(w), static_cast<uint32_t>(w >> 32));
    }
};

// Row pass fused, for products whose decompose ran the first column stages
// (DevPlan::fused_cols, N = 65536 or 32768, S = 256): the workgroup's 16 rows
// are the block rho0 .. rho0 + 15, and the four column stages left (distance
// S .. 8 S) mix them. Thread c first holds column c of the block
// (element c + S i in register i) for A and for B, runs those stages on both
// with shared twiddles, and passes them through the rows' exchange buffer into
// the phase-A layout of row_item (one workgroup barrier after each write, and
// one before the buffer is written again); row_item then runs as in
// rows_fused. The loads are 1 KB contiguous per instruction.
template <bool kGen, int kMode>
__device__ __forceinline__ void rows_fused_mid(uint32_t* __restrict__ A, const uint32_t* __restrict__ Bm,
                                               const DevPlan& P) {
    extern __shared__ uint32_t lds[];
    constexpr uint32_t M = 256, GA = 16, ROW = M + M / 16;
    const uint32_t c = threadIdx.x, t = c & 15, rl = c >> 4;
    const uint32_t rho0 = blockIdx.x << 4;
    const int j = blockIdx.y;
    const size_t base = static_cast<size_t>(j) * P.N + static_cast<size_t>(rho0) * M;
    uint32_t a[16], b[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
        a[i] = A[base + i * M + c];
        b[i] = Bm[base + i * M + c];
    }
    const uint32_t R = 1u << P.logR;
    const MidTw mt{as_const(reinterpret_cast<const uint64_t*>(P.ctw) + static_cast<size_t>(j) * R), R + rho0};
    const uint32_t p = P.pc[j].p;
    phase_b<true, true, kMode>(a, b, mt, 0, p);
    uint32_t* la = lds + rl * ROW;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) lds[i * ROW + padded(c)] = a[i];
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) a[k] = la[padded(t + 16 * k)];
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) lds[i * ROW + padded(c)] = b[i];
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) b[k] = la[padded(t + 16 * k)];
    uint2* twa = reinterpret_cast<uint2*>(lds + 16 * ROW) + rl * GA;
    row_item<8, kGen, kMode>(A, P, j, rho0 + rl, base + static_cast<size_t>(rl) * M, t, la, twa, a, b);
}

// A batch: blockIdx.z = product. Four waves per SIMD: the build schedules this
// file for instruction-level parallelism, and at up to 128 VGPRs the butterflies
// of A and B interleave (faster than seven waves of a tighter schedule).
template <bool kGen, int kMode>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_ntt_rows_fused_mid_batch(DevPlan P, BatchArgs ba) {
    if (static_cast<int>(blockIdx.y) >= ba.K[blockIdx.z]) return;   // primes past this product's K
    rows_fused_mid<kGen, kMode>(ba.A[blockIdx.z], ba.B[blockIdx.z], P);
}

bool reg_ok(uint32_t logm) { return logm >= 4 && logm <= 8; }

// Columns per column-pass workgroup: up to 512 threads, at most S columns.
uint32_t col_logC(const DevPlan& P) {
    const uint32_t want = 13 - P.logR;   // 2^(logC + logR - 4) <= 512 threads
    return want < P.logS ? want : P.logS;
}

// Rows per row-pass workgroup: up to 256 threads, at most R rows.
uint32_t row_logG(const DevPlan& P) {
    const uint32_t want = 12 - P.logS;
    return want < P.logR ? want : P.logR;
}

template <int LOGM, int LOGC, int kMode>
void col_launch3(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st, dim3 grid, dim3 block, size_t lds,
                 const BatchArgs* ba) {
    if (ba) {
        if (forward)
            launch((k_ntt_col_batch<LOGM, true, LOGC, kMode>), grid, block, lds, st, P, *ba);
        else
            launch((k_ntt_col_batch<LOGM, false, LOGC, kMode>), grid, block, lds, st, P, *ba);
    } else if (forward) {
        launch((k_ntt_col_reg<LOGM, true, LOGC, kMode>), grid, block, lds, st, X, P);
    } else {
        launch((k_ntt_col_reg<LOGM, false, LOGC, kMode>), grid, block, lds, st, X, P);
    }
}

template <int LOGM, int LOGC>
void col_launch2(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st, const BatchArgs* ba = nullptr,
                 int count = 0) {
    const dim3 grid(1u << (P.logS - LOGC), P.K, ba ? (forward ? 2 * count : count) : 1);
    const dim3 block(1u << (LOGC + LOGM - 4));
    const size_t lds = sizeof(uint32_t) * ((1u << LOGM) + (1u << (LOGM - 4))) << LOGC;
    if (P.lazy == 1)
        col_launch3<LOGM, LOGC, 1>(P, X, forward, st, grid, block, lds, ba);
    else
        col_launch3<LOGM, LOGC, 0>(P, X, forward, st, grid, block, lds, ba);
}

// (R, C) shape of the column pass as one switch key, or -1 if only the generic kernels apply.
int col_shape(const DevPlan& P) {
    if (P.logR < 4 || P.logR > 8) return -1;
    const int key = static_cast<int>(P.logR * 16 + col_logC(P));
    switch (key) {
        case 4 * 16 + 4: case 4 * 16 + 5: case 5 * 16 + 5: case 5 * 16 + 6:
        case 6 * 16 + 6: case 6 * 16 + 7: case 7 * 16 + 6: case 8 * 16 + 5: return key;
        default: return -1;
    }
}

// The (R, C) shapes that N = 2^8 .. 2^17 produce; any other goes generic.
template <int LOGM>
void col_launch(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    const uint32_t logC = col_logC(P);
    switch (LOGM * 16 + logC) {
        case 4 * 16 + 4: col_launch2<4, 4>(P, X, forward, st); return;
        case 4 * 16 + 5: col_launch2<4, 5>(P, X, forward, st); return;
        case 5 * 16 + 5: col_launch2<5, 5>(P, X, forward, st); return;
        case 5 * 16 + 6: col_launch2<5, 6>(P, X, forward, st); return;
        case 6 * 16 + 6: col_launch2<6, 6>(P, X, forward, st); return;
        case 6 * 16 + 7: col_launch2<6, 7>(P, X, forward, st); return;
        case 7 * 16 + 6: col_launch2<7, 6>(P, X, forward, st); return;
        case 8 * 16 + 5: col_launch2<8, 5>(P, X, forward, st); return;
        default: launch_ntt_cols_generic(P, X, forward, st);
    }
}

// The row pass reading its twiddles from the full tables needs them present.
void need_full_tables(const DevPlan& P) {
    if (!P.row_gen && (!P.tw || !P.itw)) throw std::logic_error("row pass from tables without the full tables");
}

template <int LOGM, int kMode>
void row_launch3(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st, dim3 grid, dim3 block, size_t lds,
                 uint32_t logG) {
    if (forward && P.row_gen)
        launch((k_ntt_row_reg<LOGM, true, true, kMode>), grid, block, lds, st, X, P, logG);
    else if (forward)
        launch((k_ntt_row_reg<LOGM, true, false, kMode>), grid, block, lds, st, X, P, logG);
    else if (P.row_gen)
        launch((k_ntt_row_reg<LOGM, false, true, kMode>), grid, block, lds, st, X, P, logG);
    else
        launch((k_ntt_row_reg<LOGM, false, false, kMode>), grid, block, lds, st, X, P, logG);
}

template <int LOGM>
void row_launch(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    need_full_tables(P);
    const uint32_t logG = row_logG(P);
    const dim3 grid(1u << (P.logR - logG), P.K);
    const dim3 block(1u << (logG + LOGM - 4));
    const size_t lds = (sizeof(uint32_t) * ((1u << LOGM) + (1u << (LOGM - 4))) << logG) +
                       (sizeof(uint2) << (LOGM - 4 + logG));
    if (P.lazy == 1)
        row_launch3<LOGM, 1>(P, X, forward, st, grid, block, lds, logG);
    else
        row_launch3<LOGM, 0>(P, X, forward, st, grid, block, lds, logG);
}

template <int LOGM>
void fused_launch_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st) {
    need_full_tables(P);
    const uint32_t logG = row_logG(P);
    const dim3 grid(1u << (P.logR - logG), P.K, count);
    const dim3 block(1u << (logG + LOGM - 4));
    const size_t lds = (sizeof(uint32_t) * ((1u << LOGM) + (1u << (LOGM - 4))) << logG) +
                       (sizeof(uint2) << (LOGM - 4 + logG));
    switch (P.lazy * 2 + (P.row_gen ? 1 : 0)) {
        case 3: launch((k_ntt_rows_fused_batch<LOGM, true, 1>), grid, block, lds, st, P, ba, logG); break;
        case 2: launch((k_ntt_rows_fused_batch<LOGM, false, 1>), grid, block, lds, st, P, ba, logG); break;
        case 1: launch((k_ntt_rows_fused_batch<LOGM, true, 0>), grid, block, lds, st, P, ba, logG); break;
        default: launch((k_ntt_rows_fused_batch<LOGM, false, 0>), grid, block, lds, st, P, ba, logG); break;
    }
}

template <int LOGM>
void fused_launch(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st) {
    need_full_tables(P);
    const uint32_t logG = row_logG(P);
    const dim3 grid(1u << (P.logR - logG), P.K);
    const dim3 block(1u << (logG + LOGM - 4));
    const size_t lds = (sizeof(uint32_t) * ((1u << LOGM) + (1u << (LOGM - 4))) << logG) +
                       (sizeof(uint2) << (LOGM - 4 + logG));
    switch (P.lazy * 2 + (P.row_gen ? 1 : 0)) {
        case 3: launch((k_ntt_rows_fused_reg<LOGM, true, 1>), grid, block, lds, st, A, B, P, logG); break;
        case 2: launch((k_ntt_rows_fused_reg<LOGM, false, 1>), grid, block, lds, st, A, B, P, logG); break;
        case 1: launch((k_ntt_rows_fused_reg<LOGM, true, 0>), grid, block, lds, st, A, B, P, logG); break;
        default: launch((k_ntt_rows_fused_reg<LOGM, false, 0>), grid, block, lds, st, A, B, P, logG); break;
    }
}

}  // namespace

void launch_ntt_cols(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    if (P.logR == 0) return;
    if (P.col_mfma) {
        launch_ntt_cols_mfma(P, X, forward, st);
        return;
    }
    switch (P.logR) {
        case 4: col_launch<4>(P, X, forward, st); return;
        case 5: col_launch<5>(P, X, forward, st); return;
        case 6: col_launch<6>(P, X, forward, st); return;
        case 7: col_launch<7>(P, X, forward, st); return;
        case 8: col_launch<8>(P, X, forward, st); return;
        default: launch_ntt_cols_generic(P, X, forward, st);
    }
}

void launch_ntt_rows(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st) {
    switch (reg_ok(P.logS) ? P.logS : 0) {
        case 4: row_launch<4>(P, X, forward, st); return;
        case 5: row_launch<5>(P, X, forward, st); return;
        case 6: row_launch<6>(P, X, forward, st); return;
        case 7: row_launch<7>(P, X, forward, st); return;
        case 8: row_launch<8>(P, X, forward, st); return;
        default: launch_ntt_rows_generic(P, X, forward, st);
    }
}

void launch_ntt_rows_fused(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st) {
    switch (reg_ok(P.logS) ? P.logS : 0) {
        case 4: fused_launch<4>(P, A, B, st); return;
        case 5: fused_launch<5>(P, A, B, st); return;
        case 6: fused_launch<6>(P, A, B, st); return;
        case 7: fused_launch<7>(P, A, B, st); return;
        case 8: fused_launch<8>(P, A, B, st); return;
        default: launch_ntt_rows_fused_generic(P, A, B, st);
    }
}

bool batch_kernels_ok(const DevPlan& P) {
    return P.logR > 0 && col_shape(P) >= 0 && reg_ok(P.logS);
}

void launch_ntt_cols_batch(const DevPlan& P, const BatchArgs& ba, int count, bool forward, hipStream_t st) {
    if (P.col_mfma) {
        launch_ntt_cols_mfma(P, nullptr, forward, st, &ba, count);
        return;
    }
    switch (col_shape(P)) {
        case 4 * 16 + 4: col_launch2<4, 4>(P, nullptr, forward, st, &ba, count); return;
        case 4 * 16 + 5: col_launch2<4, 5>(P, nullptr, forward, st, &ba, count); return;
        case 5 * 16 + 5: col_launch2<5, 5>(P, nullptr, forward, st, &ba, count); return;
        case 5 * 16 + 6: col_launch2<5, 6>(P, nullptr, forward, st, &ba, count); return;
        case 6 * 16 + 6: col_launch2<6, 6>(P, nullptr, forward, st, &ba, count); return;
        case 6 * 16 + 7: col_launch2<6, 7>(P, nullptr, forward, st, &ba, count); return;
        case 7 * 16 + 6: col_launch2<7, 6>(P, nullptr, forward, st, &ba, count); return;
        case 8 * 16 + 5: col_launch2<8, 5>(P, nullptr, forward, st, &ba, count); return;
        default: throw std::runtime_error("no batched column pass for this geometry");
    }
}

// N = 65536 (R = 256: 16 outer rows) or 32768 (R = 128: 8), rows of S = 256.
bool fused_cols_ok(const DevPlan& P) {
    return (P.logN == 16 || P.logN == 15) && P.logS == 8 && P.logR == P.logN - 8 && P.dec_digits && P.ctw &&
           (P.row_gen || (P.tw && P.itw));
}

void launch_ntt_rows_fused_mid_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st) {
    if (!fused_cols_ok(P)) throw std::logic_error("fused column stages need N = 32768 or 65536 and the matrix-core tables");
    const dim3 grid(1u << (P.logR - 4), P.K, count);   // R rows in blocks of 16
    const size_t lds = sizeof(uint32_t) * (16 * (256 + 16)) + sizeof(uint2) * 16 * 16;
    switch (P.lazy * 2 + (P.row_gen ? 1 : 0)) {
        case 3: launch((k_ntt_rows_fused_mid_batch<true, 1>), grid, dim3(256), lds, st, P, ba); break;
        case 2: launch((k_ntt_rows_fused_mid_batch<false, 1>), grid, dim3(256), lds, st, P, ba); break;
        case 1: launch((k_ntt_rows_fused_mid_batch<true, 0>), grid, dim3(256), lds, st, P, ba); break;
        default: launch((k_ntt_rows_fused_mid_batch<false, 0>), grid, dim3(256), lds, st, P, ba); break;
    }
}

void launch_ntt_rows_fused_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st) {
    switch (reg_ok(P.logS) ? P.logS : 0) {
        case 4: fused_launch_batch<4>(P, ba, count, st); return;
        case 5: fused_launch_batch<5>(P, ba, count, st); return;
        case 6: fused_launch_batch<6>(P, ba, count, st); return;
        case 7: fused_launch_batch<7>(P, ba, count, st); return;
        case 8: fused_launch_batch<8>(P, ba, count, st); return;
        default: throw std::runtime_error("no batched row pass for this geometry");
    }
}

}  // namespace pm
