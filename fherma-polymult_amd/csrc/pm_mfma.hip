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
//   x_n = sum_t u_{n,t} 256^t over its 4L bytes; T_{j,t} = 256^t mod p_j, centred
//   into (-p/2, p/2] and split into 4 balanced digits e_{j,t,b}:
//     G_b[j][n] = sum_t e_{j,t,b} (u_{n,t} - 128)        |G_b| <= 4L 2^14 <= 2^22
//     x_n mod p_j = (sum_b 256^b G_b + 128 sum_t T_{j,t}) mod p_j
//   MFMA rows are (prime, digit) pairs, columns coefficients, K-dim bytes t. The
//   four G_b of one (j, n) land in one lane and are combined in FP64 (exact below
//   2^53) and reduced with one FP64 quotient estimate.
//
// CRT (reference rns_kernels.cuh:265-364: x = sum_j y_j M_j + negvP_v mod q):
//   y_{j,n} = sum_a u_{j,n,a} 256^a (4 bytes), M_j = sum_s m_{j,s} 256^s (4L + 1
//   balanced digits). Then 256^a M_j has digits m_{j,s-a}, so
//     G[s][n] = sum_{(j,a)} m_{j,s-a} (u_{j,n,a} - 128)   |G| <= 4K 2^14 < 2^24
//     sum_j y_j M_j = sum_s 256^s G[s][n] + 0x80808080 sum_j M_j   (exactly)
//   MFMA rows are byte positions s, columns coefficients, K-dim (prime, byte).
//   A lane's four rows are the four bytes of one 32-bit limb, so it folds them
//   into one signed 64-bit limb partial; the carry chain, + negvP_v and the
//   reduction mod q run one coefficient per lane after an LDS transpose.
//   v = round(sum_j y_j / p_j) is summed in FP64 while the y bytes are loaded.
//
// MFMA operand layout (16x16x64 i8): lane l supplies row (A) / column (B)
// l % 16 and the 16 K-elements 16 (l / 16) + [0, 16) of each 64-wide K step;
// the result D[i][j] sits in lane 16 (i / 4) + j, element i % 4. The K order
// inside a lane only has to agree between A and B, which it does by
// construction.
#include "pm_bigint.h"
#include "pm_device.h"
#include "pm_mfma_host.h"

#include <stdexcept>
#include <vector>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;          // waves per workgroup
constexpr int kTiles = 4;          // 16-coefficient tiles per wave: 64 coefficients
constexpr int kDecMaxKSteps = 4;   // 4L <= 256 bytes
constexpr int kCrtMaxKSteps = 10;  // 4K <= 640

__device__ __forceinline__ v4i mfma16(const v4i& a, const v4i& b, const v4i& c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// r mod p for an integer-valued double |r| < 2^52: one FP64 quotient estimate
// (off by at most one), an exact FMA remainder, then corrections.
__device__ __forceinline__ uint32_t mod_p_f64(double r, uint32_t p, double pinv) {
    const double pd = static_cast<double>(p);
    const double qd = floor(r * pinv);
    long long ri = static_cast<long long>(fma(-qd, pd, r));
    ri += ri < 0 ? static_cast<long long>(p) : 0;
    ri -= ri >= static_cast<long long>(p) ? static_cast<long long>(p) : 0;
    return static_cast<uint32_t>(ri);
}

// Decompose, one workgroup per 64 coefficients (4 tiles of 16). Their limbs are
// copied to LDS with coalesced loads (rows of 16 KT + 4 words, zero beyond L,
// stored as int8 digits u ^ 0x80), every wave takes its B fragments from there,
// and the four waves split the prime groups. A fragments of the next group are
// loaded before the current group's MFMAs.
template <int KT>
__global__ void __launch_bounds__(64 * kWaves)
k_decompose_mfma(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res, DevPlan P, uint32_t k0,
                 uint32_t count) {
    extern __shared__ uint32_t xs[];
    constexpr int Ls = 16 * KT + 4;
    const int L = P.L;
    // per-prime constants for the epilogue: p, corr, 1/p
    double* kpinv = reinterpret_cast<double*>(xs + 16 * kTiles * Ls);
    uint32_t* kp = reinterpret_cast<uint32_t*>(kpinv + P.K);
    uint32_t* kcorr = kp + P.K;
    for (int j = threadIdx.x; j < P.K; j += blockDim.x) {
        kpinv[j] = P.p_recip[j];
        kp[j] = P.pc[j].p;
        kcorr[j] = P.dec_corr[j];
    }
    const uint32_t nb = k0 + blockIdx.x * (16 * kTiles);
    const uint32_t nk = min(static_cast<uint32_t>(16 * kTiles), k0 + count - nb);
    const uint32_t* src = limbs + static_cast<size_t>(nb) * L;
    constexpr uint32_t kSlots = 16u * kTiles * Ls;
    if (nk == 16u * kTiles) {
        // A full tile: its 64 L words are contiguous and 16-byte aligned (nb is a
        // multiple of 64). At most KT 16-byte loads per thread, all in flight.
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
                    const uint32_t i = 4 * i4 + e, row = i / L;
                    xs[row * Ls + (i - row * L)] = w4[e] ^ 0x80808080u;
                }
            }
        }
        // digits of the padding words past L: byte 0
        const uint32_t pad = Ls - L;
        for (uint32_t i = threadIdx.x; i < 16u * kTiles * pad; i += blockDim.x) {
            const uint32_t row = i / pad;
            xs[row * Ls + L + (i - row * pad)] = 0x80808080u;
        }
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
    for (int jg = wave; jg < groups; jg += kWaves) {
        v4i a[KT];
#pragma unroll
        for (int ks = 0; ks < KT; ++ks) a[ks] = an[ks];
        const int jn = min(jg + kWaves, groups - 1);
#pragma unroll
        for (int ks = 0; ks < KT; ++ks) an[ks] = dig[(jn * KT + ks) * 64 + lane];
        v4i acc[kTiles];
#pragma unroll
        for (int t = 0; t < kTiles; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KT; ++ks)
#pragma unroll
            for (int t = 0; t < kTiles; ++t) acc[t] = mfma16(a[ks], bf[t][ks], acc[t]);
        // Lane (g, cl) holds rows 4g .. 4g + 3 = prime 4 jg + g, digits b = 0..3.
        const int j = 4 * jg + g;
        if (j < P.K) {
            const uint32_t p = kp[j];
            const double pinv = kpinv[j];
            const double corr = static_cast<double>(kcorr[j]);
            uint32_t* out = res + static_cast<size_t>(j) * P.N + nb;
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                const double r = fma(16777216.0, static_cast<double>(acc[t][3]),
                                     fma(65536.0, static_cast<double>(acc[t][2]),
                                         fma(256.0, static_cast<double>(acc[t][1]),
                                             static_cast<double>(acc[t][0]) + corr)));
                const uint32_t v = mod_p_f64(r, p, pinv);
                if (static_cast<uint32_t>(16 * t + cl) < nk) out[16 * t + cl] = v;
            }
        }
    }
}

// CRT in two kernels.
//
// k_crt_mfma_partials, one workgroup per 64 coefficients:
//   1. all threads: y[j][nb .. nb + 64) for every prime into LDS as int8 digits
//      (rows of 68 words, rows past K zero), with the partial sums of y_j / p_j
//      and of y_j M_j / q, which wave 0 folds and writes out per coefficient;
//   2. the four waves split the limb tiles st: MFMA over all (prime, byte)
//      K-steps; each lane folds its four rows (the bytes of one 32-bit limb)
//      into a signed 64-bit limb partial, stored limb-major [limb][n].
// k_crt_finish, one thread per coefficient: v and the quotient estimate Q, then
//   R = sum of partials + offset correction + negvP_v - Q q in one carry chain
//   (R in [-q, 2q) since Q is exact or one off), then one correction by +-q.
constexpr int kYRow = 68;   // 64 coefficients + 4: lane groups g land on alternate bank halves

__global__ void __launch_bounds__(64 * kWaves)
k_crt_mfma_partials(const uint32_t* __restrict__ Y, long long* __restrict__ part, double2* __restrict__ fq_out,
                    DevPlan P, uint32_t k0, uint32_t count) {
    extern __shared__ uint32_t smem[];
    const int K = P.K, KS = P.crt_ksteps, ST = P.crt_stiles;
    const int KR = 16 * KS;
    uint32_t* ydig = smem;
    double* fr = reinterpret_cast<double*>(smem + static_cast<size_t>(KR) * kYRow);   // 4 x 64
    double* fq = fr + 4 * 64;                                                          // 4 x 64
    const uint32_t nb = k0 + blockIdx.x * 64;
    const uint32_t nk = min(64u, k0 + count - nb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    {
        const uint32_t col = nb + min(static_cast<uint32_t>(lane), nk - 1);
        const bool lane_in = static_cast<uint32_t>(lane) < nk;
        double f = 0, f2 = 0;
        // rows j = wave + 4 i: every load of the thread in flight at once
        constexpr int kRows = 4 * kCrtMaxKSteps;   // per wave, for KR <= 160
        uint32_t y[kRows];
#pragma unroll
        for (int i = 0; i < kRows; ++i) y[i] = Y[static_cast<size_t>(min(wave + 4 * i, K - 1)) * P.N + col];
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            const int j = wave + 4 * i;
            if (j < KR) {
                const uint32_t yv = (j < K && lane_in) ? y[i] : 0u;
                ydig[j * kYRow + lane] = yv ^ 0x80808080u;
                f = fma(static_cast<double>(yv), P.p_recip[min(j, K - 1)], f);
                f2 = fma(static_cast<double>(yv), P.mq[min(j, K - 1)], f2);
            }
        }
        fr[wave * 64 + lane] = f;
        fq[wave * 64 + lane] = f2;
    }
    __syncthreads();
    if (wave == 0 && static_cast<uint32_t>(lane) < nk) {
        fq_out[nb + lane] = make_double2(fr[lane] + fr[64 + lane] + fr[128 + lane] + fr[192 + lane],
                                         fq[lane] + fq[64 + lane] + fq[128 + lane] + fq[192 + lane]);
    }
    const int g = lane >> 4, cl = lane & 15;
    const v4i* dig = reinterpret_cast<const v4i*>(P.crt_digits);
    for (int st = wave; st < ST; st += kWaves) {
        v4i acc[kTiles];
#pragma unroll
        for (int t = 0; t < kTiles; ++t) acc[t] = v4i{0, 0, 0, 0};
        // every A fragment of this limb tile at once (clamped, not branched)
        v4i af[kCrtMaxKSteps];
#pragma unroll
        for (int ks = 0; ks < kCrtMaxKSteps; ++ks) af[ks] = dig[(st * KS + min(ks, KS - 1)) * 64 + lane];
#pragma unroll
        for (int ks = 0; ks < kCrtMaxKSteps; ++ks) {
            if (ks >= KS) continue;
            const uint32_t* yr = ydig + (16 * ks + 4 * g) * kYRow + cl;
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                v4i b;
                b[0] = static_cast<int>(yr[16 * t]);
                b[1] = static_cast<int>(yr[kYRow + 16 * t]);
                b[2] = static_cast<int>(yr[2 * kYRow + 16 * t]);
                b[3] = static_cast<int>(yr[3 * kYRow + 16 * t]);
                acc[t] = mfma16(af[ks], b, acc[t]);
            }
        }
        // rows 16 st + 4 g + r, r = 0..3: the bytes of limb 4 st + g
        long long* out = part + static_cast<size_t>(4 * st + g) * P.N + nb + cl;
#pragma unroll
        for (int t = 0; t < kTiles; ++t) {
            if (static_cast<uint32_t>(16 * t + cl) < nk) {
                out[16 * t] = static_cast<long long>(acc[t][0]) + static_cast<long long>(acc[t][1]) * 256 +
                              static_cast<long long>(acc[t][2]) * 65536 +
                              static_cast<long long>(acc[t][3]) * 16777216;
            }
        }
    }
}

template <int LMAX>
__global__ void __launch_bounds__(256)
k_crt_finish(const long long* __restrict__ part, const double2* __restrict__ fq_in, uint32_t* __restrict__ c,
             DevPlan P, uint32_t k0, uint32_t count) {
    const uint32_t n = k0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= k0 + count) return;
    const int K = P.K, L = P.L, NL = 4 * P.crt_stiles;   // limbs with partials
    const double2 fv = fq_in[n];
    const int v = crt_round(fv.x, K);
    const uint64_t Q = quotient_estimate(fv.y + P.nq[v]);
    const uint32_t Ql = static_cast<uint32_t>(Q), Qh = static_cast<uint32_t>(Q >> 32);
    const uint32_t* nv = P.negvP + static_cast<size_t>(v) * L;
    uint32_t x[LMAX + 2];
#pragma unroll
    for (int l = 0; l < LMAX; ++l) x[l] = nv[min(l, L - 1)];   // all in flight, clamped
    long long carry = 0;
    uint64_t a1 = 0, b1 = 0, b2 = 0;
    const long long* pp = part + n;
#pragma unroll
    for (int l0 = 0; l0 < LMAX + 2; l0 += 8) {
        long long pv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) pv[u] = pp[static_cast<size_t>(min(l0 + u, NL - 1)) * P.N];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int l = l0 + u;
            if (l < LMAX + 2) {
                const uint32_t qv = l < L ? P.q[l] : 0u;
                long long s = carry + (l < L + 2 ? static_cast<long long>(P.crt_corr[l]) : 0) -
                              q_times_limb(Ql, Qh, qv, a1, b1, b2);
                if (l < L) s += x[l];
                if (l < NL) s += pv[u];
                x[l] = static_cast<uint32_t>(s);
                carry = s >> 32;   // arithmetic: partials may be negative
            }
        }
    }
    fix_range<LMAX>(x, P.q, L);
    uint32_t* out = c + static_cast<size_t>(n) * L;
    if ((L & 3) == 0 && (reinterpret_cast<uintptr_t>(c) & 15) == 0) {
        // 16-byte stores: a quarter of the cache-line accesses of word stores
        uint4* o4 = reinterpret_cast<uint4*>(out);
#pragma unroll
        for (int l = 0; l < LMAX; l += 4)
            if (l < L) o4[l / 4] = make_uint4(x[l], x[l + 1], x[l + 2], x[l + 3]);
    } else {
#pragma unroll
        for (int l = 0; l < LMAX; ++l)
            if (l < L) out[l] = x[l];
    }
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
    if (L > 4 * kDecMaxKSteps * 4 || (4 * K + 63) / 64 > kCrtMaxKSteps) {
        throw std::runtime_error("point too large for the matrix-core tables");
    }
    // Decompose tables.
    const int KT = (4 * L + 63) / 64;
    const int KG = (K + 3) / 4;
    std::vector<int8_t> dec(static_cast<size_t>(KG) * KT * 64 * 16, 0);
    std::vector<uint32_t> dcorr(K);
    for (int j = 0; j < K; ++j) {
        const uint64_t p = h.p[j];
        uint64_t T = 1, sum = 0;
        for (int t = 0; t < 4 * L; ++t) {
            sum = (sum + T) % p;
            const long long centred = T <= p / 2 ? static_cast<long long>(T) : static_cast<long long>(T) - static_cast<long long>(p);
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
        dcorr[j] = static_cast<uint32_t>(sum * 128 % p);
    }
    d.dec_digits = static_cast<const int8_t*>(upload(dec.data(), dec.size()));
    d.dec_corr = static_cast<const uint32_t*>(upload(dcorr.data(), dcorr.size() * sizeof(uint32_t)));
    d.dec_ksteps = KT;

    // CRT tables.
    const int KS = (4 * K + 63) / 64;
    const int ST = (4 * L + 4 + 15) / 16;
    const int nd = 4 * L + 1;
    std::vector<int8_t> m(static_cast<size_t>(K) * nd);
    for (int j = 0; j < K; ++j) balanced_digits(h.M.data() + static_cast<size_t>(j) * L, L, m.data() + static_cast<size_t>(j) * nd, nd);
    std::vector<int8_t> crt(static_cast<size_t>(ST) * KS * 64 * 16, 0);
    for (int st = 0; st < ST; ++st)
        for (int ks = 0; ks < KS; ++ks)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 16; ++e) {
                    const int s = 16 * st + lane % 16;
                    const int j = 16 * ks + 4 * (lane / 16) + e / 4;
                    const int a = e % 4;
                    const int idx = s - a;
                    if (j < K && idx >= 0 && idx < nd)
                        crt[((static_cast<size_t>(st) * KS + ks) * 64 + lane) * 16 + e] = m[static_cast<size_t>(j) * nd + idx];
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
    d.crt_digits = static_cast<const int8_t*>(upload(crt.data(), crt.size()));
    d.crt_corr = static_cast<const uint32_t*>(upload(ccorr.data(), ccorr.size() * sizeof(uint32_t)));
    d.crt_ksteps = KS;
    d.crt_stiles = ST;
}

void launch_decompose_mfma(const DevPlan& P, const uint32_t* limbs, uint32_t* res, hipStream_t st) {
    launch_decompose_mfma_range(P, limbs, res, 0, P.N, st);
}

void launch_decompose_mfma_range(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0,
                                 uint32_t count, hipStream_t st) {
    if (!P.dec_digits) throw std::runtime_error("matrix-core tables not prepared");
    if (count == 0) return;
    const dim3 grid((count + 16 * kTiles - 1) / (16 * kTiles));
    const size_t lds = sizeof(uint32_t) * 16 * kTiles * (16 * P.dec_ksteps + 4) + 16 * static_cast<size_t>(P.K);
    switch (P.dec_ksteps) {
        case 1: hipLaunchKernelGGL(k_decompose_mfma<1>, grid, dim3(64 * kWaves), lds, st, limbs, res, P, k0, count); break;
        case 2: hipLaunchKernelGGL(k_decompose_mfma<2>, grid, dim3(64 * kWaves), lds, st, limbs, res, P, k0, count); break;
        case 3: hipLaunchKernelGGL(k_decompose_mfma<3>, grid, dim3(64 * kWaves), lds, st, limbs, res, P, k0, count); break;
        case 4: hipLaunchKernelGGL(k_decompose_mfma<4>, grid, dim3(64 * kWaves), lds, st, limbs, res, P, k0, count); break;
        default: throw std::runtime_error("L out of range");
    }
}

void launch_crt_mfma(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st) {
    launch_crt_mfma_range(P, Y, c, 0, P.N, st);
}

void launch_crt_mfma_range(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count,
                           hipStream_t st) {
    if (!P.crt_digits || !P.crt_part) throw std::runtime_error("matrix-core tables not prepared");
    if (count == 0) return;
    const dim3 grid((count + 63) / 64);
    const size_t lds = sizeof(uint32_t) * (static_cast<size_t>(16 * P.crt_ksteps) * kYRow + 4 * 64 * 2 * 2);
    hipLaunchKernelGGL(k_crt_mfma_partials, grid, dim3(64 * kWaves), lds, st, Y, P.crt_part, P.crt_fq, P, k0,
                       count);
    const dim3 grid2((count + 255) / 256);
    const int lmax = (P.L + 7) / 8 * 8;
    switch (lmax) {
#define PM_CRT_CASE(LM) \
    case LM: hipLaunchKernelGGL(k_crt_finish<LM>, grid2, dim3(256), 0, st, P.crt_part, P.crt_fq, c, P, k0, count); break;
        PM_CRT_CASE(8)
        PM_CRT_CASE(16)
        PM_CRT_CASE(24)
        PM_CRT_CASE(32)
        PM_CRT_CASE(40)
        PM_CRT_CASE(48)
        PM_CRT_CASE(56)
        PM_CRT_CASE(64)
#undef PM_CRT_CASE
        default: throw std::runtime_error("L out of range");
    }
}

}  // namespace pm
