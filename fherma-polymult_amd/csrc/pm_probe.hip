// Issue-rate probes for gfx950: how many cycles a SIMD spends per wave64
// instruction of the kinds the kernels are built from. No counterpart in the
// reference; the bench reports these figures beside the kernels' instruction
// counts (a kernel's VALU issue time is instructions x waves x cycles / SIMDs).
//
// Every probe runs independent dependency chains (8 VALU chains, or 4 MFMA
// accumulators) so that latency does not hide the issue rate, with 16 waves
// per CU, and writes one word per thread so that nothing is optimized away.
//   kind 0  v_add_u32            5  v_permlane32_swap_b32
//        1  v_mul_lo_u32         6  v_permlane16_swap_b32
//        2  v_mul_hi_u32         7  v_mfma_i32_16x16x64_i8
//        3  v_mad_u64_u32        8  MFMA waves and v_mul_lo_u32 waves on the same SIMDs
//        4  v_mad_i64_i32        9  one wave stream alternating MFMA and v_mul_lo_u32
//                               10  MFMA waves and v_add_u32 waves on the same SIMDs
//                               11  one wave stream alternating MFMA and v_add_u32
#include "pm_state.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <vector>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kChains = 8, kUnroll = 8;   // 64 VALU instructions per iteration
constexpr int kMfmaChains = 4, kMfmaUnroll = 4;   // 16 MFMAs per iteration

// 64-bit multiply-adds: each chain alternates two accumulators (the
// destination may not overlap a source), and the new high word is the next
// multiplier, so no product is loop-invariant.
template <bool kSigned>
__device__ __forceinline__ void mad64_iter(uint32_t (&x)[kChains], uint64_t (&acc)[kChains], uint32_t y) {
    uint64_t alt[kChains];
#pragma unroll
    for (int u = 0; u < kUnroll; u += 2) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if constexpr (kSigned)
                alt[k] = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(x[k])) * static_cast<int32_t>(y) +
                                               static_cast<int64_t>(acc[k]));
            else
                alt[k] = static_cast<uint64_t>(x[k]) * y + acc[k];
            asm volatile("" : "+v"(alt[k]));
            x[k] = static_cast<uint32_t>(alt[k] >> 32);
            if constexpr (kSigned)
                acc[k] = static_cast<uint64_t>(static_cast<int64_t>(static_cast<int32_t>(x[k])) * static_cast<int32_t>(y) +
                                               static_cast<int64_t>(alt[k]));
            else
                acc[k] = static_cast<uint64_t>(x[k]) * y + alt[k];
            asm volatile("" : "+v"(acc[k]));
            x[k] = static_cast<uint32_t>(acc[k] >> 32);
        }
    }
}

template <int KIND>
__device__ __forceinline__ void valu_iter(uint32_t (&x)[kChains], uint64_t (&acc)[kChains], uint32_t y) {
    if constexpr (KIND == 3 || KIND == 4) {
        mad64_iter<KIND == 4>(x, acc, y);
        return;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if constexpr (KIND == 0) {
                x[k] = x[k] + y;
            } else if constexpr (KIND == 1) {
                x[k] = x[k] * y;
            } else if constexpr (KIND == 2) {
                x[k] = __umulhi(x[k], y);
            } else if constexpr (KIND == 5) {
                if (k % 2 == 0) {
                    const auto r = __builtin_amdgcn_permlane32_swap(x[k], x[k + 1], false, false);
                    x[k] = r[0];
                    x[k + 1] = r[1];
                }
            } else if constexpr (KIND == 6) {
                if (k % 2 == 0) {
                    const auto r = __builtin_amdgcn_permlane16_swap(x[k], x[k + 1], false, false);
                    x[k] = r[0];
                    x[k + 1] = r[1];
                }
            }
            // opaque to the optimizer (no instruction): repeated operations stay
            // as many instructions instead of being folded into one
            asm volatile("" : "+v"(x[k]));
        }
    }
}

__device__ __forceinline__ void mfma_iter(v4i (&acc)[kMfmaChains], const v4i& a, const v4i& b) {
#pragma unroll
    for (int u = 0; u < kMfmaUnroll; ++u)
#pragma unroll
        for (int k = 0; k < kMfmaChains; ++k) {
            acc[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc[k], 0, 0, 0);
            asm volatile("" : "+v"(acc[k]));
        }
}

template <int KIND>
__global__ void __launch_bounds__(512) k_probe(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                               int iters) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = in[threadIdx.x & 63] | 1u;
    uint32_t x[kChains];
    uint64_t acc[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
        x[k] = in[(threadIdx.x + 7 * k) & 63] + t;
        acc[k] = x[k];
    }
    v4i m[kMfmaChains];
    const v4i a = {static_cast<int>(x[0]), static_cast<int>(x[1]), static_cast<int>(x[2]), static_cast<int>(x[3])};
    const v4i b = {static_cast<int>(x[4]), static_cast<int>(x[5]), static_cast<int>(x[6]), static_cast<int>(x[7])};
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k) m[k] = v4i{k, 0, 0, 0};
    const bool mfma_wave = KIND == 7 || ((KIND == 8 || KIND == 10) && (threadIdx.x >> 6) < 4);
    for (int it = 0; it < iters; ++it) {
        if constexpr (KIND <= 6) {
            valu_iter<KIND>(x, acc, y);
        } else if constexpr (KIND == 7) {
            mfma_iter(m, a, b);
        } else if constexpr (KIND == 8 || KIND == 10) {
            if (mfma_wave)
                mfma_iter(m, a, b);
            else
                valu_iter<KIND == 8 ? 1 : 0>(x, acc, y);
        } else {
            mfma_iter(m, a, b);
            valu_iter<KIND == 9 ? 1 : 0>(x, acc, y);
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += x[k] + static_cast<uint32_t>(acc[k] >> 32) + static_cast<uint32_t>(acc[k]);
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k) s += m[k][0] + m[k][1] + m[k][2] + m[k][3];
    out[t] = s;
}

template <int KIND>
void probe_launch(const uint32_t* in, uint32_t* out, int cus, int iters, hipStream_t st) {
    launch(k_probe<KIND>, dim3(2 * cus), dim3(512), 0, st, in, out, iters);
}

}  // namespace

// Runs probe `kind` and returns its wall time (ms) and the instructions each
// SIMD issued: VALU instructions, MFMAs (kinds 7-11 count both).
void probe_issue(int kind, int iters, double* ms, double* valu_per_simd, double* mfma_per_simd, double* clock_mhz) {
    int dev = 0, cus = 0, khz = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    check(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev), "hipDeviceGetAttribute");
    const size_t threads = static_cast<size_t>(2 * cus) * 512;
    uint32_t* in = static_cast<uint32_t*>(dev_alloc(64 * sizeof(uint32_t)));
    uint32_t* out = static_cast<uint32_t*>(dev_alloc(threads * sizeof(uint32_t)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(in);
        (void)hipFree(out);
    };
    try {
        std::vector<uint32_t> h(64);
        for (int i = 0; i < 64; ++i) h[i] = 0x9E3779B9u * (i + 1);
        check(hipMemcpy(in, h.data(), 64 * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        auto run = [&](int n) {
            switch (kind) {
                case 0: probe_launch<0>(in, out, cus, n, nullptr); break;
                case 1: probe_launch<1>(in, out, cus, n, nullptr); break;
                case 2: probe_launch<2>(in, out, cus, n, nullptr); break;
                case 3: probe_launch<3>(in, out, cus, n, nullptr); break;
                case 4: probe_launch<4>(in, out, cus, n, nullptr); break;
                case 5: probe_launch<5>(in, out, cus, n, nullptr); break;
                case 6: probe_launch<6>(in, out, cus, n, nullptr); break;
                case 7: probe_launch<7>(in, out, cus, n, nullptr); break;
                case 8: probe_launch<8>(in, out, cus, n, nullptr); break;
                case 9: probe_launch<9>(in, out, cus, n, nullptr); break;
                case 10: probe_launch<10>(in, out, cus, n, nullptr); break;
                case 11: probe_launch<11>(in, out, cus, n, nullptr); break;
                default: throw std::invalid_argument("no such probe");
            }
        };
        run(1);   // warm-up
        check(hipEventRecord(e0, nullptr), "hipEventRecord");
        run(iters);
        check(hipEventRecord(e1, nullptr), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0;
        check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        // 16 waves per CU = 4 per SIMD; kinds 8 and 10 put 2 MFMA and 2 VALU waves on each SIMD
        const double waves = 4.0, valu_inst = kChains * kUnroll, mfma_inst = kMfmaChains * kMfmaUnroll;
        const bool pair = kind == 5 || kind == 6;   // one swap instruction serves two chains
        *ms = t;
        *valu_per_simd = kind <= 6   ? waves * iters * valu_inst / (pair ? 2 : 1)
                         : kind == 8 || kind == 10 ? 2.0 * iters * valu_inst
                         : kind == 9 || kind == 11 ? waves * iters * valu_inst
                                                   : 0.0;
        *mfma_per_simd = kind == 7 || kind == 9 || kind == 11 ? waves * iters * mfma_inst
                         : kind == 8 || kind == 10             ? 2.0 * iters * mfma_inst
                                                               : 0.0;
        *clock_mhz = khz / 1000.0;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace pm
