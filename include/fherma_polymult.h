/*
 * fherma_polymult.h -- C ABI of the MI355X-native exact negacyclic product
 *     c = a * b mod (X^N + 1, q)
 * with N a power of two up to 2^22 and q of up to 2048 bits (L <= 64 u32 limbs).
 *
 * Plain pointers and sizes only. Operands are N x L little-endian u32 limbs,
 * coefficient-major (the spec's wire layout, reference fherma.h:34-41); every
 * coefficient of a and b must be below q. The output is canonical in [0, q).
 *
 * Every function returns PM_OK (0) or a non-zero pm_status; pm_last_error()
 * then holds a message (per thread). A plan is not re-entrant: calls on one plan
 * must not overlap (the reference's state is not either, solve.cu:871).
 *
 * The C++ entry points of the FHERMA boundary (include/fherma.h) are thin
 * wrappers over pm_plan_create / pm_multiply / pm_plan_destroy.
 */
#ifndef FHERMA_POLYMULT_H
#define FHERMA_POLYMULT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_plan pm_plan;

enum pm_status {
    PM_OK = 0,
    PM_EINVAL = 1,    /* invalid point or argument (the reference throws here) */
    PM_EHIP = 2,      /* a HIP runtime call failed */
    PM_ENOMEM = 3,    /* device or host allocation failed */
    PM_EINTERNAL = 4
};

enum pm_form {
    PM_FORM_PLAIN = 0, /* VALU decompose and CRT */
    PM_FORM_MFMA = 1   /* exact int8 matrix-core decompose and CRT (default) */
};

/* Last error message of the calling thread ("" if none). */
const char* pm_last_error(void);

/* --- plan: fherma_init (reference solve.cu:710-869, rns_plan.h:118-249) ----
 * Builds the prime set with the reference's rule (primes p < 2^31,
 * p = 1 mod max(2N, 2^16), downward, until sum log2 p > 2 bits(q) + log2 N + 21),
 * every table, and the device workspace. Refuses N outside [2, 2^22] or not a
 * power of two, L outside [1, 64], q whose top limb is zero, and points that
 * need more primes than exist. W is recorded, not checked (as in the reference). */
int pm_plan_create(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, pm_plan** out);

/* fherma_free (reference solve.cu:1061-1083). Null is accepted. */
void pm_plan_destroy(pm_plan* plan);

/* N, L and K of a plan; any output pointer may be null. */
int pm_plan_query(const pm_plan* plan, uint32_t* N, uint32_t* L, uint32_t* K);

/* The plan's K primes, in the reference's order (rns_plan.h:144-153). */
int pm_plan_primes(const pm_plan* plan, uint32_t* out, uint32_t cap);

/* The prime rule alone, on the host (no device): the K primes for a point with
 * N coefficients and a q of `qbits` bits (reference rns_plan.h:138-155). */
int pm_choose_primes(uint32_t N, uint32_t qbits, uint32_t* out, uint32_t cap, uint32_t* K);

/* Selects the plain or the matrix-core form for later calls on this plan. */
int pm_plan_set_form(pm_plan* plan, int form);

/* Row-pass twiddle source (reference rns_plan.h:269-280): the full K x N
 * tables, or generation from per-row and per-prime factors. Chosen at plan
 * creation by the reference's rule (generate when 16 K N bytes exceed 64 MB)
 * unless FHERMA_TWIDDLES=table|gen|auto says otherwise (any other value makes
 * plan creation fail with PM_EINVAL); both give identical results. */
typedef enum {
    PM_TWIDDLES_TABLE = 0,
    PM_TWIDDLES_GEN = 1
} pm_twiddles;
int pm_plan_set_twiddles(pm_plan* plan, int source);
int pm_plan_get_twiddles(const pm_plan* plan);   /* pm_twiddles, or -1 for a null plan */

/* The NTT's column passes (the stages of butterfly distance >= S, reference
 * solve.cu:128-166): butterflies on the VALU, or, for R = 256 rows (N = 65536),
 * the exact int8 matrix-core form (two radix-16 steps as v_mfma_i32_16x16x64_i8
 * products per column, pm_ntt_mfma.hip). Both give the same NTT-domain values
 * mod p. The default is the VALU form (PM_NTT_FORM=valu or auto);
 * PM_NTT_FORM=mfma makes new plans use the matrix-core form where it applies
 * (any other value makes plan creation fail with PM_EINVAL). No counterpart in
 * the reference, which has only butterflies. */
typedef enum {
    PM_NTT_VALU = 0,
    PM_NTT_MFMA = 1
} pm_ntt_form;
int pm_plan_set_ntt_form(pm_plan* plan, int form);
int pm_plan_get_ntt_form(const pm_plan* plan);   /* pm_ntt_form, or -1 for a null plan */

/* Fused column stages (N = 65536 or 32768, matrix-core form): 1 runs the
 * product's first forward column stages (four or three: R/16-point columns of
 * stride 16 N/R) inside the decompose kernel and the next four (those within a
 * block of 16 rows) at the start of the fused row pass, so the forward column
 * passes, and their write and re-read of both K x N residue arrays, are gone; 0
 * runs the passes of the reference's schedule (solve.cu:128-233). The
 * NTT-domain values are the same.
 * New plans take 1 where the geometry allows unless PM_FUSED_COLS=0 (1 or auto:
 * the default; any other value makes plan creation fail with PM_EINVAL). Turning
 * it on where the geometry does not allow it fails with PM_EINVAL. The stage
 * entry points and the host pipeline's chunked decompose always run the
 * unfused kernels. No counterpart in the reference. */
int pm_plan_set_fused_cols(pm_plan* plan, int on);
int pm_plan_get_fused_cols(const pm_plan* plan);   /* 0 or 1, or -1 for a null plan */

/* --- products -------------------------------------------------------------
 * pm_multiply: host pointers, synchronous; the fherma_run equivalent
 * (reference solve.cu:871-1059). Inputs cross PCIe through pinned staging.
 * pm_multiply_device: device pointers, enqueued on `stream` (a hipStream_t;
 * null means the plan's own stream), returns without waiting. */
int pm_multiply(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c);
/* Products on host pointers (pm_multiply's path), up to `depth` at once, each
 * through its own plan's pipeline: one product's inputs cross PCIe while
 * another's kernels run and a third's answer comes back. Products of the same
 * plan run one after another. depth <= 0: PM_HOST_DEPTH, else 3. Synchronous.
 * No counterpart in the reference, which runs one product per fherma_run. */
int pm_batch_multiply_host(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                           uint32_t* const* c, int depth);
int pm_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream);

/* pm_multiply_device without any allocation: no counterpart in the reference.
 * It is for callers that cannot allocate device memory at call time (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same kernels run with the two K x N residue arrays in a static device
 * workspace of 72 MiB per GPU (held until the process ends), so the product
 * must need 2 K N <= 18874368 words: every point up to N = 65536 with a q of
 * up to 2048 bits; larger ones fail with PM_EINVAL before any launch. The plan
 * (created beforehand, which allocates) supplies only its read-only tables.
 * Every call on one GPU shares that workspace: calls that overlap on a GPU
 * (two streams, two threads, two plans) race and must be ordered by the
 * caller. Nothing in the workspace needs clearing between calls. */
int pm_multiply_device_noalloc(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream);

/* How pm_multiply moves host operands (reference solve.cu:14-34, 282-296):
 *   PM_HOST_STAGED  chunked copies through pinned staging (parallel host memcpy,
 *                   hipMemcpyAsync), overlapped with the decompose and CRT chunks
 *   PM_HOST_MAPPED  the decompose reads, and the CRT writes, pinned host memory
 *                   mapped into the device, with no DMA copies
 *   PM_HOST_DIRECT  hipMemcpyAsync from and to the caller's pageable memory
 * The reference's XNACK/HMM mode (kernels on pageable memory) needs xnack+ code
 * objects and HSA_XNACK=1, which this build does not assume. */
enum pm_host_mode { PM_HOST_STAGED = 0, PM_HOST_MAPPED = 1, PM_HOST_DIRECT = 2 };
int pm_plan_set_host_mode(pm_plan* plan, int mode);   /* both directions */
/* The input side (a, b) and the output side (c) separately, as the reference's
 * FHERMA_IN and FHERMA_OUT do (solve.cu:13-34). */
int pm_plan_set_host_modes(pm_plan* plan, int in_mode, int out_mode);
int pm_plan_get_host_modes(const pm_plan* plan, int* in_mode, int* out_mode);
/* Times the host modes on dummy operands and keeps the fastest; staged stays
 * unless another is at least 5 % faster (reference solve.cu:817-862). Each
 * timed call sizes a fresh answer on the helper thread as fherma_run does,
 * after a write over 4 x the operand size to clear the host caches; one
 * warm-up, median of seven. The candidates: staged and mapped both ways at 1, 2
 * and 4 chunks, mapped in / staged out, staged in / mapped out, direct. The
 * mode out is the input side's. fherma_init calls it unless FHERMA_IN,
 * FHERMA_OUT or PM_HOST_MODE names a mode or FHERMA_AUTOTUNE (PM_AUTOTUNE) is 0.
 * Either output may be null. */
int pm_plan_tune_host(pm_plan* plan, int* mode, double* ms_per_product);
/* The timed choice of the plan's row-pass twiddle source (pm_plan_set_twiddles)
 * and, where it exists, its column-pass form (pm_plan_set_ntt_form): device-
 * resident products on dummy operands, two warm-up runs and the median of five
 * per candidate; the current setting (the reference's size rule for the
 * twiddles, rns_plan.h:269-280) is kept unless another is at least 5 % faster.
 * FHERMA_TWIDDLES / PM_NTT_FORM in the environment pin their dimension.
 * fherma_init calls it after the host choice, under the same switch
 * (FHERMA_AUTOTUNE). ms_per_product, if not null, receives the chosen setting's
 * median. */
int pm_plan_tune_kernels(pm_plan* plan, double* ms_per_product);

/* A batch of independent products, each with its own plan (N and q may differ),
 * all device pointers, enqueued on `stream`. No counterpart in the reference:
 * it is for callers with many products per call (uneven batches). With an
 * explicit stream, products of distinct plans in the matrix-core form at one N
 * (any q width: their prime lists are prefixes of one list) run as batched
 * launches, one kernel per phase covering all of them (decompose and CRT: one
 * per K-step count); the rest product by product. Without a stream, product by
 * product on each plan's stream. */
int pm_batch_multiply_device(int count, pm_plan* const* plans, const uint32_t* const* a,
                             const uint32_t* const* b, uint32_t* const* c, void* stream);

/* The layout of a captured batch.
 *   runs      products that can share batched launches (matrix-core form, one
 *             N) are cut into about this many runs, each set of one N into a
 *             share by its work, as contiguous slices of equal work after
 *             sorting by K; each run is one chain of batched launches. 0: every
 *             product as its own chain of launches (no batched launches).
 *   streams   graph branches the runs (or chains) are spread over, largest
 *             work first onto the least-loaded branch.
 *   ntt_form  pm_ntt_form of the batched runs' column passes, or -1 for the
 *             plans' own.
 *   twiddles  pm_twiddles of the batched runs' row passes, or -1 for the plans'
 *             own. */
typedef struct {
    int runs;
    int streams;
    int ntt_form;
    int twiddles;
} pm_batch_shape;

/* The same batch captured once as a hipGraph and replayed by pm_batch_launch:
 * one launch per batch, no per-kernel host overhead. pm_batch_create picks the
 * shape by timing: it captures the candidate shapes (runs, streams; then the
 * column-pass form and the row-pass twiddle source of the batched runs), replays
 * each (one warm-up, the median of five) and keeps the fastest, so creation
 * runs the batch a few dozen times: a and b must hold readable device data and
 * c receives the products. No counterpart in the reference, which times only
 * its copy modes (solve.cu:817-862). The environment overrides a dimension:
 * PM_BATCH_RUNS, PM_BATCH_STREAMS, PM_NTT_FORM, FHERMA_TWIDDLES (the last two
 * as the plans' own); PM_BATCH_TUNE=0 skips the timing (runs 16, streams 16,
 * the plans' own forms). pm_batch_create_shaped captures one given shape and
 * times nothing. The device pointers and the plans' forms are fixed at
 * creation; each product needs its own plan, and the plans must outlive the
 * batch, as must every stream a batch was launched on (pm_batch_create times
 * on the null stream). A batch works in its plans' residue buffers: its
 * launches must not overlap each other or other calls on those plans. */
typedef struct pm_batch pm_batch;
int pm_batch_create(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                    uint32_t* const* c, pm_batch** out);
int pm_batch_create_shaped(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                           uint32_t* const* c, const pm_batch_shape* shape, pm_batch** out);
int pm_batch_launch(pm_batch* batch, void* stream);
/* The batch's shape as captured (streams: the branches used; ntt_form and
 * twiddles -1 where the plans' own were kept) and, when pm_batch_create timed
 * it, the median replay in ms of the chosen shape (else 0). Either output may
 * be null. */
int pm_batch_get_shape(const pm_batch* batch, pm_batch_shape* shape, double* ms);
/* 1 if any products of the batch were captured as batched launches, 0 if every
 * product as its own chain. */
int pm_batch_phased(const pm_batch* batch);
void pm_batch_destroy(pm_batch* batch);

/* --- the five steps, one at a time (device pointers, enqueued on `stream`) --
 * Residues are K x N u32, prime-major. The forward NTT leaves the merged
 * Cooley-Tukey transform's bit-reversed order (rns_kernels.cuh:159-174);
 * pointwise multiplies by N^-1 (P/p_j)^-1, so the inverse NTT ends in
 * y_j = c * (P/p_j)^-1 mod p_j; crt maps y back to N x L limbs mod q. */
int pm_stage_decompose(pm_plan* plan, const uint32_t* limbs, uint32_t* res, void* stream);
int pm_stage_ntt(pm_plan* plan, uint32_t* res, int forward, void* stream);
int pm_stage_pointwise(pm_plan* plan, uint32_t* A, const uint32_t* B, void* stream);
int pm_stage_crt(pm_plan* plan, const uint32_t* res, uint32_t* c, void* stream);

/* --- timing (HIP events on the stream the kernels run on) -------------------
 * pm_time_multiply_device: mean milliseconds per product over `iters` calls
 * after `warmup` untimed ones.
 * pm_profile_phases: mean milliseconds per launch of each of the PM_NPHASES
 * phases (named by pm_phase_name()), each phase's kernels run `iters` times back
 * to back between two events on the plan's stream; c holds the product after. */
#define PM_NPHASES 7
int pm_time_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, int warmup,
                            int iters, double* ms_per_product, void* stream);
int pm_profile_phases(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, int iters,
                      double* ms_per_phase, void* stream);
const char* pm_phase_name(int phase);

/* --- device memory and streams, for callers without an allocator -----------*/
int pm_device_alloc(size_t bytes, void** out);
int pm_device_free(void* ptr);
int pm_copy_to_device(void* dst, const void* src, size_t bytes);
int pm_copy_to_host(void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and stores),
 * enqueued on `stream`; pointers 16-byte aligned, bytes a multiple of 16. No
 * counterpart in the reference: it is the known-size yardstick against which
 * the HBM traffic counters (FETCH_SIZE, WRITE_SIZE) of the product's kernels
 * are scaled. */
int pm_device_copy(void* dst, const void* src, size_t bytes, void* stream);
/* The sum mod 2^64 of the 32-bit words of a device buffer (16-byte aligned,
 * bytes a multiple of 16), by a read-only streaming kernel; synchronous, on
 * the null stream. No counterpart in the reference: the read-only yardstick
 * for FETCH_SIZE, and a checksum of device results. */
int pm_device_sum(const void* src, size_t bytes, uint64_t* sum);
int pm_stream_create(void** out);
int pm_stream_destroy(void* stream);
int pm_stream_synchronize(void* stream);
int pm_device_synchronize(void);
/* Issue-rate probe of the current device, no counterpart in the reference: the
 * wall time (ms) of a kernel issuing only one kind of instruction from
 * independent chains at 16 waves per CU, and the instructions each SIMD issued
 * (VALU and MFMA), with the device's peak clock, from which cycles per
 * instruction follow. Kinds: 0 v_add_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32,
 * 3 v_mad_u64_u32, 4 v_mad_i64_i32, 5 v_permlane32_swap, 6 v_permlane16_swap,
 * 7 v_mfma_i32_16x16x64_i8, 8 MFMA waves beside v_mul_lo_u32 waves on each
 * SIMD, 9 one wave stream alternating MFMA and v_mul_lo_u32, 10 MFMA waves
 * beside v_add_u32 waves, 11 one wave stream alternating MFMA and v_add_u32. */
int pm_probe_issue(int kind, int iters, double* ms, double* valu_per_simd, double* mfma_per_simd, double* clock_mhz);

/* Device name and compute-unit count of the current device. */
int pm_device_info(char* name, size_t len, int* compute_units);

#ifdef __cplusplus
}
#endif

#endif /* FHERMA_POLYMULT_H */
