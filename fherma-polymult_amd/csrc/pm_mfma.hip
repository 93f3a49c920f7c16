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

__global__ void __launch_bounds__(64 * kWaves)
k_decompose_mfma(const uint32_t* __restrict__ limbs, uint32_t* __restrict__ res, DevPlan P) {
    const int lane = threadIdx.x & 63;
    const uint32_t nbase = (blockIdx.x * kWaves + (threadIdx.x >> 6)) * (16 * kTiles);
    if (nbase >= P.N) return;   // no barriers below
    const int g = lane >> 4, cl = lane & 15;
    const int L = P.L, KT = P.dec_ksteps;

    // B fragments: coefficient n = nbase + 16 t + cl, bytes 64 ks + 16 g + [0, 16),
    // i.e. limb words 16 ks + 4 g + [0, 4). Words past L pair with zero digits.
    v4i bf[kTiles][kDecMaxKSteps];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const uint32_t n = nbase + 16 * t + cl;
        const uint32_t* src = limbs + static_cast<size_t>(n) * L;
#pragma unroll
        for (int ks = 0; ks < kDecMaxKSteps; ++ks) {
            v4i v = {0, 0, 0, 0};
            if (ks < KT && n < P.N) {
                const int w0 = 16 * ks + 4 * g;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (w0 + i < L) v[i] = static_cast<int>(src[w0 + i]);
            }
            bf[t][ks] = v ^ static_cast<int>(0x80808080u);
        }
    }

    const v4i* dig = reinterpret_cast<const v4i*>(P.dec_digits);
    const int groups = (P.K + 3) >> 2;
    for (int jg = 0; jg < groups; ++jg) {
        v4i acc[kTiles];
#pragma unroll
        for (int t = 0; t < kTiles; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < kDecMaxKSteps; ++ks) {
            if (ks < KT) {
                const v4i a = dig[(jg * KT + ks) * 64 + lane];
#pragma unroll
                for (int t = 0; t < kTiles; ++t) acc[t] = mfma16(a, bf[t][ks], acc[t]);
            }
        }
        // Lane (g, cl) holds rows 4g .. 4g + 3 = prime 4 jg + g, digits b = 0..3.
        const int j = 4 * jg + g;
        if (j < P.K) {
            const uint32_t p = P.pc[j].p;
            const double pinv = P.p_recip[j];
            const double corr = static_cast<double>(P.dec_corr[j]);
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                const uint32_t n = nbase + 16 * t + cl;
                if (n < P.N) {
                    const double r = fma(16777216.0, static_cast<double>(acc[t][3]),
                                         fma(65536.0, static_cast<double>(acc[t][2]),
                                             fma(256.0, static_cast<double>(acc[t][1]),
                                                 static_cast<double>(acc[t][0]) + corr)));
                    res[static_cast<size_t>(j) * P.N + n] = mod_p_f64(r, p, pinv);
                }
            }
        }
    }
}

template <int LMAX>
__global__ void __launch_bounds__(64 * kWaves)
k_crt_mfma(const uint32_t* __restrict__ Y, uint32_t* __restrict__ c, DevPlan P) {
    __shared__ long long part[kWaves][4][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const uint32_t nbase = (blockIdx.x * kWaves + wave) * (16 * kTiles);
    const int K = P.K, L = P.L, KS = P.crt_ksteps, ST = P.crt_stiles;

    // B fragments: coefficient n = nbase + 16 t + cl, K-elements (prime, byte)
    // for primes 16 ks + 4 g + [0, 4): the four words y[j][n] as they are.
    v4i bf[kTiles][kCrtMaxKSteps];
    double f[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const uint32_t n = nbase + 16 * t + cl;
        f[t] = 0;
#pragma unroll
        for (int ks = 0; ks < kCrtMaxKSteps; ++ks) {
            v4i v = {0, 0, 0, 0};
            if (ks < KS) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int j = 16 * ks + 4 * g + i;
                    if (j < K && n < P.N) {
                        const uint32_t y = Y[static_cast<size_t>(j) * P.N + n];
                        f[t] = fma(static_cast<double>(y), P.p_recip[j], f[t]);
                        v[i] = static_cast<int>(y);
                    }
                }
            }
            bf[t][ks] = v ^ static_cast<int>(0x80808080u);
        }
    }
    // sum_j y_j / p_j: add the four lane groups' partial sums.
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        f[t] += __shfl_xor(f[t], 16);
        f[t] += __shfl_xor(f[t], 32);
    }
    // From here on lane l owns coefficient nbase + l, i.e. tile g, column cl.
    const double frac = g == 0 ? f[0] : g == 1 ? f[1] : g == 2 ? f[2] : f[3];
    const int v = crt_round(frac, K);
    const uint32_t* nv = P.negvP + static_cast<size_t>(v) * L;
    const v4i* dig = reinterpret_cast<const v4i*>(P.crt_digits);

    uint32_t x[LMAX + 2];
    long long carry = 0;
    constexpr int kSTiles = LMAX / 4 + 1;   // covers limbs 0 .. LMAX + 3
#pragma unroll
    for (int st = 0; st < kSTiles; ++st) {
        if (st < ST) {
            v4i acc[kTiles];
#pragma unroll
            for (int t = 0; t < kTiles; ++t) acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < kCrtMaxKSteps; ++ks) {
                if (ks < KS) {
                    const v4i a = dig[(st * KS + ks) * 64 + lane];
#pragma unroll
                    for (int t = 0; t < kTiles; ++t) acc[t] = mfma16(a, bf[t][ks], acc[t]);
                }
            }
            // rows 16 st + 4 g + r, r = 0..3: the bytes of limb 4 st + g.
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                part[wave][g][16 * t + cl] = static_cast<long long>(acc[t][0]) +
                                             static_cast<long long>(acc[t][1]) * 256 +
                                             static_cast<long long>(acc[t][2]) * 65536 +
                                             static_cast<long long>(acc[t][3]) * 16777216;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lam = 4 * st + i;
            if (lam < LMAX + 2) {
                long long s = carry;
                if (st < ST) s += part[wave][i][lane];
                if (lam < L + 2) s += P.crt_corr[lam];
                if (lam < L) s += nv[lam];
                x[lam] = static_cast<uint32_t>(s);
                carry = s >> 32;   // arithmetic: partials may be negative, the total is not
            }
        }
        __syncthreads();
    }
    reduce_mod_q<LMAX>(x, P);
    const uint32_t n = nbase + lane;
    if (n < P.N) {
        uint32_t* out = c + static_cast<size_t>(n) * L;
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
    if (!P.dec_digits) throw std::runtime_error("matrix-core tables not prepared");
    const uint32_t per_block = 16 * kTiles * kWaves;
    const dim3 grid((P.N + per_block - 1) / per_block);
    hipLaunchKernelGGL(k_decompose_mfma, grid, dim3(64 * kWaves), 0, st, limbs, res, P);
}

void launch_crt_mfma(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st) {
    if (!P.crt_digits) throw std::runtime_error("matrix-core tables not prepared");
    const uint32_t per_block = 16 * kTiles * kWaves;
    const dim3 grid((P.N + per_block - 1) / per_block);
    const int lmax = (P.L + 7) / 8 * 8;
    switch (lmax) {
#define PM_CRT_CASE(LM) \
    case LM: hipLaunchKernelGGL(k_crt_mfma<LM>, grid, dim3(64 * kWaves), 0, st, Y, c, P); break;
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
