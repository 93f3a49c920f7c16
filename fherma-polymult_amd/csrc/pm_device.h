// Device-side view of a plan, and the kernel launchers (pm_device.hip,
// pm_mfma.hip). Residues are K x N u32, prime-major; limbs are N x L u32,
// coefficient-major with little-endian limbs (the spec's wire layout).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm {

// Per-prime constants, 32 bytes, read with scalar loads.
struct PrimeConst {
    uint32_t p;       // the prime, p < 2^31, p = 1 mod max(2N, 2^16)
    uint32_t pneg;    // -p^-1 mod 2^32
    uint32_t c32;     // 2^32 mod p
    uint32_t c32s;    // its Shoup companion
    uint32_t c64;     // 2^64 mod p
    uint32_t one_s;   // Shoup companion of 1: floor(2^32 / p)
    uint32_t scale;   // N^-1 (P/p_j)^-1 2^32 mod p
    uint32_t scale_s; // its Shoup companion
};

struct DevPlan {
    uint32_t N, logN, logS, logR;
    int L, K;
    int row_gen;              // row-pass twiddles: 1 generated (rz, rom), 0 from the full tables
    // Butterfly form of the register NTT kernels (pm_ntt.hip), used when every
    // pass has one: 1 forward residues in [0, 2p); 0 canonical (the generic
    // kernels, or a geometry without register kernels for every pass).
    int lazy;
    const PrimeConst* pc;     // K
    const uint32_t* pow2;     // K x L: 2^(32 l) mod p_j
    // The b side of a product: tables that decompose b into b scale_j mod p_j
    // (scale_j the pointwise constant of PrimeConst), so that the fused row
    // pass needs only a Montgomery product. pow2b: K x L; the MFMA tables
    // dec_digits_b / dec_init_b are laid out as dec_digits / dec_init.
    const uint32_t* pow2b;
    const int8_t* dec_digits_b;
    const int2* dec_init_b;
    // Column pass on the matrix cores (pm_ntt_mfma.hip), R = 256: per prime and
    // direction, 17 matrices as MFMA fragments (87040 bytes); null when absent.
    const uint8_t* colm_fwd;
    const uint8_t* colm_inv;
    int col_mfma;             // 1: the column passes run on the matrix cores (tables present)
    // 1: a product's forward NTT runs its first column stages (R/16-point
    // columns of stride 16 N/R) inside the matrix-core decompose and the next
    // four (those within a block of 16 rows) at the start of the fused row
    // pass (launch_decompose_cols_batch, launch_ntt_rows_fused_mid_batch):
    // N = 65536 or 32768 (S = 256), matrix-core form. Read only by the product
    // paths that launch those kernels; every other launcher ignores it.
    int fused_cols;
    // Full tables, null unless the row pass reads its twiddles from tables or
    // the geometry needs the generic kernels.
    const uint2* tw;          // K x N: (psi^bitrev(i), Shoup companion)
    const uint2* itw;         // K x N: (psi^-bitrev(i) 2^32 mod p, Montgomery companion)
    const uint2* ctw;         // K x R: the first R pairs of each prime's tw (the column pass's)
    const uint2* citw;        // K x R: the same of itw
    // Row-pass twiddles, generated: the twiddle of row rho, stage distance 2^e,
    // block b is Z[rho][e] * Om[b] (pm_ntt.hip).
    const uint2* rz;          // K x R x logS: (Z 2^32 mod p, Shoup companion), forward
    const uint2* rzi;         // inverse
    const uint32_t* rom;      // K x S/2: Om, forward
    const uint32_t* romi;     // inverse
    const double* p_recip;    // K
    const uint32_t* M;        // K x L: P/p_j mod q
    const uint32_t* negvP;    // (K + 1) x L: -vP mod q
    const uint32_t* q;        // L
    // Quotient estimate for the final reduction: floor(T / q) with
    // T = sum_j y_j M_j + negvP_v is sum_j y_j mq[j] + nq[v] in FP64, exact or
    // one off (the accumulated error stays below 2^-6).
    const double* mq;         // K: M_j / q
    const double* nq;         // K + 1: negvP_v / q
    // int8 matrix-core forms (pm_mfma.hip); null until prepared.
    const int8_t* dec_digits; // [ceil(K/4)][dec_ksteps][64 lanes][16]: MFMA A fragments
    // K: the initial values of prime j's digit-0 and digit-2 accumulators
    // (x, y): with them the folded sum R = G_0 + 256 G_1 + 2^16 (G_2 + 256 G_3)
    // is positive, below 2^48, and congruent to the exact sum x T (T the table
    // entries 256^t 2^32 mod p_j times the side's factor; pm_mfma.hip)
    const int2* dec_init;
    // CRT digit table, column-major: word J kTabRows + R (pm_mfma.hip) holds
    // byte R - 3 of the balanced digits of M_j for primes j = 4J .. 4J + 3
    // (zero for j >= K and outside digit rows 0 .. 4L).
    const uint32_t* crt_tab;
    // CRT constants, zero-padded so that every index the kernel forms is in
    // range: crt_fc = 1 / p_j (160 entries), then M_j / q (160); crt_lc = q,
    // -P mod q and the offset correction, 72 words each, limb l at l + 2.
    const double* crt_fc;
    const uint32_t* crt_lc;
    const uint32_t* crt_corr; // L + 2 limbs: 128 (2^32 - 1)/255 sum_j M_j, exact
    int dec_ksteps;           // ceil(4L / 64)
    int crt_ksteps;           // ceil(4K / 64)
    int crt_stiles;           // ceil((4L + 4) / 16)
    uint32_t dec_magic;       // floor(2^32 / D) + 1, D = L / 4 if 4 | L and L >= 8, else L: i / D = mulhi(i, dec_magic) for i < 2^12
};

// P with the b-side decompose tables in place of the plain ones.
inline DevPlan b_side(const DevPlan& P) {
    DevPlan Q = P;
    Q.pow2 = P.pow2b;
    Q.dec_digits = P.dec_digits_b;
    Q.dec_init = P.dec_init_b;
    return Q;
}

// Every kernel of the library is launched through this one function.
template <typename F, typename... Args>
inline void launch(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st, Args... args) {
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

// Plain form (VALU): every launcher enqueues on `st` and returns.
void launch_decompose_valu(const DevPlan& P, const uint32_t* limbs, uint32_t* res, hipStream_t st);
void launch_ntt_cols(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
void launch_ntt_rows(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
// Forward rows of A and B, pointwise product into A, inverse rows; B must come
// from the b-side decompose (b_side).
void launch_ntt_rows_fused(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st);
void launch_pointwise(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st);
void launch_crt_valu(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st);
void launch_decompose_valu_range(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0,
                                 uint32_t count, hipStream_t st);
void launch_crt_valu_range(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count,
                           hipStream_t st);

// The same NTT passes with one LDS barrier per stage, for any sub-transform
// length up to 4096 (pm_device.hip); the launchers above use them when the
// register kernels of pm_ntt.hip (lengths 16 .. 256) do not apply.
void launch_ntt_cols_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
void launch_ntt_rows_generic(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st);
void launch_ntt_rows_fused_generic(const DevPlan& P, uint32_t* A, const uint32_t* B, hipStream_t st);

// Matrix-core form (int8 MFMA, exact). lazy_out: residues may be left in
// [0, 2p) (for the forward NTT of a product when P.lazy is set), else canonical.
void launch_decompose_mfma(const DevPlan& P, const uint32_t* limbs, uint32_t* res, hipStream_t st,
                           bool lazy_out = false);
// The tables of one product that depend on its K, L or q, as its batched
// launches read them (one per plan, in device memory, immutable once the plan
// exists; the kernels read it with scalar loads).
struct ProductDesc {
    int K, L;
    int dec_ksteps, crt_ksteps;
    uint32_t dec_magic, pad;
    const int8_t* dec_digits;
    const int2* dec_init;
    const int8_t* dec_digits_b;
    const int2* dec_init_b;
    const uint32_t* crt_tab;
    const double* crt_fc;
    const double* nq;
    const uint32_t* crt_lc;
};

// Batches of up to kBatchMax products that share N and whose prime lists are
// prefixes of one list (the prime rule walks one sequence per N, so products
// at one N differ only in how many primes they take): one plan P, that of the
// product with the most primes, supplies the geometry, the per-prime constants
// and the twiddles; per product, its operands a, b, answer c, residue buffers
// A, B, its prime count K and its own tables d. Each batched launch covers
// every product at once (blockIdx.z picks the product; workgroups of primes
// past a product's K return at once) and is large enough to run in steady
// state. Decompose and CRT launches take products of one K-step count.
constexpr int kBatchMax = 32;
struct BatchArgs {
    const uint32_t* a[kBatchMax];
    const uint32_t* b[kBatchMax];
    uint32_t* c[kBatchMax];
    uint32_t* A[kBatchMax];
    uint32_t* B[kBatchMax];
    const ProductDesc* d[kBatchMax];
    int K[kBatchMax];
};

// A pointer into the constant address space: loads at uniform addresses become
// scalar loads (the kernels never write what it points to).
template <class T>
using kconst_ptr = const T __attribute__((address_space(4)))*;
template <class T>
__device__ __forceinline__ kconst_ptr<T> as_const(const T* p) {
    return (kconst_ptr<T>)p;
}

bool batch_kernels_ok(const DevPlan& P);   // register NTT kernels exist for this geometry
// The column pass on the matrix cores, one launch (ba: a batch of `count`
// products, as launch_ntt_cols_batch; null: X alone).
void launch_ntt_cols_mfma(const DevPlan& P, uint32_t* X, bool forward, hipStream_t st, const BatchArgs* ba = nullptr,
                          int count = 0);
void launch_decompose_mfma_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st);
void launch_ntt_cols_batch(const DevPlan& P, const BatchArgs& ba, int count, bool forward, hipStream_t st);
void launch_ntt_rows_fused_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st);
void launch_crt_mfma_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st);
// fused_cols plans (DevPlan::fused_cols): the decompose of a and b (side -1),
// of a only (0) or of b only (1), each followed by the first four forward
// column stages, and the fused row pass that starts with the next four.
bool fused_cols_ok(const DevPlan& P);   // the geometry and tables these kernels need
void launch_decompose_cols_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st, int side = -1);
void launch_ntt_rows_fused_mid_batch(const DevPlan& P, const BatchArgs& ba, int count, hipStream_t st);

// Plain grid-stride copy of `bytes` (a multiple of 16) with 16-byte accesses.
void launch_copy16(void* dst, const void* src, size_t bytes, hipStream_t st);
// Read-only pass over `bytes` (a multiple of 16): out[0] = the sum mod 2^64 of
// its 32-bit words; partials holds kSumBlocks words of workspace.
constexpr unsigned kSumBlocks = 8192;
void launch_sum16(uint64_t* partials, uint64_t* out, const void* src, size_t bytes, hipStream_t st);

void launch_crt_mfma(const DevPlan& P, const uint32_t* Y, uint32_t* c, hipStream_t st);
// The same over coefficients [k0, k0 + count) only (chunked host pipeline).
void launch_decompose_mfma_range(const DevPlan& P, const uint32_t* limbs, uint32_t* res, uint32_t k0,
                                 uint32_t count, hipStream_t st, bool lazy_out = false);
void launch_crt_mfma_range(const DevPlan& P, const uint32_t* Y, uint32_t* c, uint32_t k0, uint32_t count,
                           hipStream_t st);

}  // namespace pm
