// The NTT's column pass on the int8 matrix cores (gfx950), for R = 256 rows.
//
// The column pass runs the stages of the merged Cooley-Tukey NTT whose butterfly
// distance is at least S (reference solve.cu:128-166, rns_kernels.cuh:128-174).
// On one column (R elements at stride S) these 8 stages are one 256-point
// twisted transform: with zeta = psi^S, a primitive 512-th root of unity,
//     Y[k] = sum_i x[i] zeta^((2k + 1) i),   written at row bitrev_8(k),
// and the inverse stages (Gentleman-Sande, psi^-1) compute
//     x'[i] = sum_k Y[k] zeta^(-(2k + 1) i) = R x[i]   from Y[k] at row bitrev_8(k).
// Each is split radix-16 x 16 (i = a + 16 b, k = c + 16 d) into two steps with
// one 16 x 16 matrix each and a twiddle between them, all depending only on the
// prime, so every column of a prime shares them:
//   forward  step 1  U[c][a]   = zeta^(a (2c + 1)) sum_b zeta^(16 b (2c + 1)) x[a + 16 b]
//            step 2  Y[c+16d]  = sum_a zeta^(32 a d) U[c][a]
//   inverse  step 1  V[a][c]   = zeta^(-(2c + 1) a) sum_d zeta^(-32 d a) Y[c + 16 d]
//            step 2  x'[a+16b] = sum_c zeta^(-16 b (2c + 1)) V[a][c]
// A step is four v_mfma_i32_16x16x64_i8 products: the K dimension holds the 16
// summed values x 4 bytes of each (an unsigned byte u enters as u - 128), the
// matrix entries are W 256^byte 2^32 mod p, centred and split into 4 balanced
// base-256 digits, one product per digit. For one output the four int32 sums G_0..G_3
// satisfy |G| <= 64 * 128 * 128 = 2^20; with the byte bias and a positive offset
// folded into the accumulators' initial values (digit 0: c0, |c0| < p/2; digit 3:
// 2^21), R = sum 256^b G_b lies in (0, 2^46) and is congruent to the exact sum times
// 2^32, so one Montgomery step t = (R + m p) / 2^32, m = R (-p^-1) mod 2^32, gives the
// output in [0, 2^31 + 2^14), below 2p for every prime above 2^30 + 2^14: any 32-bit
// value is a valid input to the next step, the forward pass leaves [0, 2p) (the
// row pass's input range) and the inverse pass ends with one conditional
// subtraction (canonical, as the CRT needs). The twiddle is one Montgomery
// product with a pair (w' = w 2^32 mod p, w' (-p^-1) mod 2^32): for any 32-bit t,
// (t w' + m p) / 2^32 with m = t w' (-p^-1) mod 2^32 is congruent to t w and below
// 2p (mul_lo and two v_mad_u64_u32).
//
// A wave owns 16 adjacent columns of one prime (two waves per workgroup, so that
// a workgroup's row segments are whole 128-byte lines). Step 1 runs 16 rounds (one
// per batch value: a forward, c inverse) with the data as the B operand (lane
// (g, cl): column cl, summed values 4g .. 4g + 3). Step 2 needs, per round, the 16
// step-1 outputs of one value of the other index in the same layout; they are
// spread over the four 16-lane groups, and 16 transposes of 4 x 4 blocks (lane
// group x register) with v_permlane32_swap / v_permlane16_swap bring them there
// without LDS. Both matrices' fragments stay in registers; the 256 twiddle pairs
// of the prime are in LDS (2 KB per workgroup, the kernel's only LDS).
#include "pm_device.h"
#include "pm_mfma_host.h"

#include "pm_arith.h"

#include <cstring>
#include <stdexcept>
#include <vector>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kFrags = 5;                     // four digit fragments and the offset fragment
constexpr uint32_t kMatBytes = kFrags * 64 * 16;
constexpr uint32_t kTwBytes = 256 * 8;        // twiddle pairs [round][output]
constexpr uint32_t kPrimeBytes = 2 * kMatBytes + kTwBytes;   // 12288 bytes per prime and direction
constexpr int kWaves = 2;                     // waves per workgroup: 32 adjacent columns, whole 128-byte lines
constexpr uint32_t kLdsBytes = kTwBytes;

__device__ __forceinline__ v4i mfma16(const v4i& a, const v4i& b, const v4i& c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t brev4(uint32_t x) { return __brev(x) >> 28; }

// R = G0 + 2^8 G1 + 2^16 G2 + 2^24 G3 (the offsets are in G0 and G3, so R is in
// (0, 2^46)), then R 2^-32 mod p by one Montgomery step, left in [0, 2^31 + 2^14).
// k16 = 2^16, kept opaque to the compiler so that R is one v_mad_i64_i32 rather
// than two 64-bit shifts and adds: six VALU instructions per output in all.
__device__ __forceinline__ uint32_t fold4(int G0, int G1, int G2, int G3, uint32_t p, uint32_t pneg, int k16) {
    const int lo = G0 + G1 * 256;
    const int hi = G2 + G3 * 256;
    const int64_t R = static_cast<int64_t>(hi) * k16 + static_cast<int64_t>(lo);
    const uint32_t m = static_cast<uint32_t>(R) * pneg;
    const uint64_t T = static_cast<uint64_t>(m) * p + static_cast<uint64_t>(R);
    return static_cast<uint32_t>(T >> 32);
}

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
    __device__ __forceinline__ v4i ld4(uint32_t voff, uint32_t soff) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        return v4i{static_cast<int>(v[0]), static_cast<int>(v[1]), static_cast<int>(v[2]), static_cast<int>(v[3])};
    }
};

// One matrix's fragments for this lane: four digits and the offsets (as the
// initial value of digit 0's accumulator).
struct Frag {
    v4i d[4], c0;
};

__device__ __forceinline__ Frag load_frag(const Buf& tab, uint32_t lane16, uint32_t mat) {
    Frag f;
#pragma unroll
    for (int k = 0; k < 4; ++k) f.d[k] = tab.ld4(lane16, mat * kMatBytes + k * 1024);
    f.c0 = tab.ld4(lane16, mat * kMatBytes + 4 * 1024);
    return f;
}

// One step on one B fragment: the four digit products and the fold.
__device__ __forceinline__ void step(const Frag& f, const v4i& b, uint32_t p, uint32_t pneg, int k16,
                                     uint32_t (&out)[4]) {
    const v4i zero = {0, 0, 0, 0}, lift = {1 << 21, 1 << 21, 1 << 21, 1 << 21};
    const v4i g0 = mfma16(f.d[0], b, f.c0);
    const v4i g1 = mfma16(f.d[1], b, zero);
    const v4i g2 = mfma16(f.d[2], b, zero);
    const v4i g3 = mfma16(f.d[3], b, lift);
#pragma unroll
    for (int s = 0; s < 4; ++s) out[s] = fold4(g0[s], g1[s], g2[s], g3[s], p, pneg, k16);
}

// t w 2^32 2^-32 mod p, below 2p, for any 32-bit t and a twiddle pair (w', w' (-p^-1)).
__device__ __forceinline__ uint32_t twiddle(uint32_t t, uint32_t wm, uint32_t wq, uint32_t p) {
    const uint32_t m = t * wq;
    const uint64_t T = static_cast<uint64_t>(m) * p + static_cast<uint64_t>(t) * wm;
    return static_cast<uint32_t>(T >> 32);
}

// Registers r[h] hold, in lane group g (lanes 16g .. 16g + 15), entry M[g][h] of a
// 4 x 4 block; afterwards r[k] holds M[k][g] there. permlane32_swap(x, y) returns
// ([x lanes 0-31, y lanes 0-31], [x lanes 32-63, y lanes 32-63]); permlane16_swap
// does the same for the 16-lane rows (0, 2 | 1, 3) of each half.
__device__ __forceinline__ void transpose4(uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3) {
    const auto a = __builtin_amdgcn_permlane32_swap(r0, r2, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(r1, r3, false, false);
    const auto c = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
    const auto d = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
    r0 = c[0];
    r1 = c[1];
    r2 = d[0];
    r3 = d[1];
}

__device__ __forceinline__ v4i biased(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    return v4i{static_cast<int>(w0 ^ 0x80808080u), static_cast<int>(w1 ^ 0x80808080u),
               static_cast<int>(w2 ^ 0x80808080u), static_cast<int>(w3 ^ 0x80808080u)};
}

template <bool kFwd>
__device__ __forceinline__ void col_mfma(uint32_t* __restrict__ X, const DevPlan& P, int j, uint32_t block) {
    extern __shared__ uint32_t smem[];
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, cl = lane & 15, wave = threadIdx.x >> 6;
    const uint32_t tile = kWaves * block + wave;
    const uint32_t S = 1u << P.logS, rowb = 4u * S;
    const Buf x(X + static_cast<size_t>(j) * P.N, 4u * P.N);
    const Buf tab((kFwd ? P.colm_fwd : P.colm_inv) + static_cast<size_t>(j) * kPrimeBytes, kPrimeBytes);
    const uint32_t col4 = 4u * (tile * 16u + cl), lane16 = 16u * lane;
    const uint32_t p = P.pc[j].p, pneg = P.pc[j].pneg;
    int k16 = 65536;
    asm volatile("" : "+s"(k16));

    // The input: round r, word k of lane (g, cl) is the summed index 4g + k of round r.
    //   forward  x[a + 16 b] at row a + 16 b: a = r, b = 4g + k
    //   inverse  Y[c + 16 d] at row bitrev_8(c + 16 d) = brev4(d) + 16 brev4(c): c = r, d = 4g + k
    // Issued round by round, so that round 0 waits only for its own four loads.
    uint32_t in[16][4], vo[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) vo[k] = col4 + rowb * (kFwd ? 16u * (4u * g + k) : brev4(4u * g + k));
#pragma unroll
    for (int r = 0; r < 16; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) in[r][k] = x.ld(vo[k], rowb * (kFwd ? r : 16u * brev4(r)));
    Frag f = load_frag(tab, lane16, 0);   // step 1's matrix, then step 2's
    // the twiddle pairs [round][output] (2 KB) to LDS
    const uint2* tw = reinterpret_cast<const uint2*>(smem);
    {
        const Buf twb((kFwd ? P.colm_fwd : P.colm_inv) + static_cast<size_t>(j) * kPrimeBytes + 2 * kMatBytes, kTwBytes);
        static_assert(64 * kWaves * 16 == kTwBytes, "one 16-byte piece per thread");
        const v4i v = twb.ld4(16u * threadIdx.x, 0);
        *reinterpret_cast<v4i*>(smem + 4u * threadIdx.x) = v;
    }
    __syncthreads();

    // Step 1 and the twiddle: y[r][s] is output 4g + s of round r.
    uint32_t y[16][4];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        step(f, biased(in[r][0], in[r][1], in[r][2], in[r][3]), p, pneg, k16, y[r]);
        const uint4 w01 = *reinterpret_cast<const uint4*>(tw + 16u * r + 4u * g);
        const uint4 w23 = *reinterpret_cast<const uint4*>(tw + 16u * r + 4u * g + 2u);
        y[r][0] = twiddle(y[r][0], w01.x, w01.y, p);
        y[r][1] = twiddle(y[r][1], w01.z, w01.w, p);
        y[r][2] = twiddle(y[r][2], w23.x, w23.y, p);
        y[r][3] = twiddle(y[r][3], w23.z, w23.w, p);
    }
    f = load_frag(tab, lane16, 1);   // in flight during the transposes
    // Step 2's round o = 4k + s needs, in lane group g, rounds 4g + t (t = 0..3) of
    // output o, held by lane group k as y[4g + t][s]: for each (s, t) the block
    // M[k][h] = y[4h + t][s] of lane group k is transposed, after which lane group
    // g holds round 4g + t of output 4k + s in y[4k + t][s].
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) transpose4(y[t][s], y[4 + t][s], y[8 + t][s], y[12 + t][s]);

    // Step 2: round o = 4k + s writes outputs e = 4g + s' (s' = 0..3).
    //   forward  Y[o + 16 e] at row brev4(e) + 16 brev4(o)
    //   inverse  x'[o + 16 e] at row o + 16 e, reduced to [0, p)
    uint32_t so[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) so[e] = col4 + rowb * (kFwd ? brev4(4u * g + e) : 16u * (4u * g + e));
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t o = 4u * k + s;
            uint32_t z[4];
            step(f, biased(y[4 * k][s], y[4 * k + 1][s], y[4 * k + 2][s], y[4 * k + 3][s]), p, pneg, k16, z);
            const uint32_t su = rowb * (kFwd ? 16u * brev4(o) : o);
#pragma unroll
            for (int e = 0; e < 4; ++e) x.st(kFwd ? z[e] : min(z[e], z[e] - p), so[e], su);
        }
}

// The forward kernel needs 172 VGPRs (two waves per SIMD), the inverse 168
// (three); capping the forward one at 168 made the compiler spill, slower.
template <bool kFwd>
__global__ void __launch_bounds__(64 * kWaves)
k_ntt_col_mfma(uint32_t* __restrict__ X, DevPlan P) {
    col_mfma<kFwd>(X, P, blockIdx.y, blockIdx.x);
}

// A batch: forward, blockIdx.z = 2 (product) + (0: A, 1: B); inverse, blockIdx.z = product (A).
template <bool kFwd>
__global__ void __launch_bounds__(64 * kWaves)
k_ntt_col_mfma_batch(DevPlan P, BatchArgs ba) {
    const int i = kFwd ? blockIdx.z >> 1 : blockIdx.z;
    if (static_cast<int>(blockIdx.y) >= ba.K[i]) return;   // primes past this product's K
    col_mfma<kFwd>(kFwd && (blockIdx.z & 1) ? ba.B[i] : ba.A[i], P, blockIdx.y, blockIdx.x);
}

// Balanced base-256 digits of |x| < 2^31.
void digits4(long long x, int8_t out[4]) {
    for (int i = 0; i < 4; ++i) {
        const long long r = ((x + 128) % 256 + 256) % 256 - 128;
        out[i] = static_cast<int8_t>(r);
        x = (x - r) / 256;
    }
    if (x != 0) throw std::logic_error("column-pass tables: value out of range");
}

long long centred(uint64_t v, uint64_t p) {
    return v <= p / 2 ? static_cast<long long>(v) : static_cast<long long>(v) - static_cast<long long>(p);
}

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return a * b % p; }

// One matrix M[row][sum] (16 x 16, entries mod p) as this kernel's fragments:
// digit fragment k, lane l = 16 g + row, byte e holds digit k of the centred
// M[row][4g + e/4] 256^(e % 4) 2^32 mod p; the offset fragment holds, for lane
// group g, c0 of rows 4g .. 4g + 3: c0 = (128 sum_K W - 2^45) mod p, centred.
void put_matrix(const std::vector<uint64_t>& M, uint64_t p, uint8_t* out) {
    const uint64_t r32 = (uint64_t(1) << 32) % p;
    long long rowsum[16] = {};
    for (int l = 0; l < 64; ++l) {
        const int g = l / 16, row = l % 16;
        for (int e = 0; e < 16; ++e) {
            const int sum = 4 * g + e / 4, byte = e % 4;
            uint64_t w = mulmod(M[row * 16 + sum], r32, p);
            for (int t = 0; t < byte; ++t) w = mulmod(w, 256, p);
            const long long c = centred(w, p);
            rowsum[row] += c;
            int8_t d[4];
            digits4(c, d);
            for (int k = 0; k < 4; ++k) out[k * 1024 + l * 16 + e] = static_cast<uint8_t>(d[k]);
        }
    }
    const long long P = static_cast<long long>(p);
    const long long lift = static_cast<long long>((uint64_t(1) << 45) % p);
    for (int l = 0; l < 64; ++l) {
        const int g = l / 16;
        for (int s = 0; s < 4; ++s) {
            const long long want = ((128 % P) * (rowsum[4 * g + s] % P) % P + P) % P;
            const int32_t c0 = static_cast<int32_t>(centred(static_cast<uint64_t>(((want - lift) % P + P) % P), p));
            std::memcpy(out + 4 * 1024 + l * 16 + 4 * s, &c0, 4);
        }
    }
}

}  // namespace

bool col_mfma_ok(const HostPlan& h) {
    if (h.logR != 8 || h.logS < 5 || h.logS > 8) return false;   // S a multiple of 32 columns
    for (uint32_t p : h.p)
        if (p <= (1u << 30) + (1u << 14)) return false;   // the fold's output must stay below 2p
    return true;
}

void build_col_mfma_tables(const HostPlan& h, bool forward, std::vector<uint8_t>& out) {
    out.assign(static_cast<size_t>(h.K) * kPrimeBytes, 0);
    std::vector<uint64_t> M(256), zp(512);
    for (int j = 0; j < h.K; ++j) {
        const uint64_t p = h.p[j];
        // zeta = psi^S (forward) or psi^-S (inverse): its 512 powers
        uint64_t z = 1, base = 1;
        const uint64_t root = forward ? h.psi[j] : h.ipsi[j];
        for (uint32_t i = 0; i < (1u << h.logS); ++i) base = mulmod(base, root, p);
        for (int e = 0; e < 512; ++e) {
            zp[e] = z;
            z = mulmod(z, base, p);
        }
        uint8_t* dst = out.data() + static_cast<size_t>(j) * kPrimeBytes;
        // step 1: rows are its outputs (c forward, a inverse), columns the summed index
        for (int row = 0; row < 16; ++row)
            for (int s = 0; s < 16; ++s)
                M[row * 16 + s] = forward ? zp[(16 * s * (2 * row + 1)) % 512] : zp[(32 * s * row) % 512];
        put_matrix(M, p, dst);
        // step 2 (d from a forward, b from c inverse)
        for (int row = 0; row < 16; ++row)
            for (int s = 0; s < 16; ++s)
                M[row * 16 + s] = forward ? zp[(32 * s * row) % 512] : zp[(16 * row * (2 * s + 1)) % 512];
        put_matrix(M, p, dst + kMatBytes);
        // twiddles [round r][step-1 output o] as Montgomery pairs:
        // forward zeta^(r (2o + 1)) (r = a, o = c), inverse zeta^-((2r + 1) o) (r = c, o = a)
        const uint32_t pinv_neg = h.pneg[j];
        uint32_t* tw = reinterpret_cast<uint32_t*>(dst + 2 * kMatBytes);
        for (int r = 0; r < 16; ++r)
            for (int o = 0; o < 16; ++o) {
                const uint64_t w = forward ? zp[(r * (2 * o + 1)) % 512] : zp[((2 * r + 1) * o) % 512];
                const uint32_t wm = static_cast<uint32_t>((w << 32) % p);
                tw[2 * (16 * r + o)] = wm;
                tw[2 * (16 * r + o) + 1] = wm * pinv_neg;
            }
    }
}

void launch_ntt_cols_mfma(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st, const BatchArgs* ba,
                          int count) {
    if (!P.colm_fwd || !P.colm_inv || P.logR != 8) throw std::runtime_error("matrix-core column pass not prepared");
    const dim3 grid((1u << P.logS) / (16 * kWaves), P.K, ba ? (forward ? 2 * count : count) : 1);
    const dim3 block(64 * kWaves);
    if (ba) {
        if (forward)
            launch(k_ntt_col_mfma_batch<true>, grid, block, kLdsBytes, st, P, *ba);
        else
            launch(k_ntt_col_mfma_batch<false>, grid, block, kLdsBytes, st, P, *ba);
    } else if (forward) {
        launch(k_ntt_col_mfma<true>, grid, block, kLdsBytes, st, X, P);
    } else {
        launch(k_ntt_col_mfma<false>, grid, block, kLdsBytes, st, X, P);
    }
}

}  // namespace pm
