// Internal state shared by the library's host-side files: the plan object
// behind the opaque pm_plan handle, and the HIP error helpers.
#pragma once

#include "../../include/fherma_polymult.h"
#include "pm_device.h"
#include "pm_mfma_host.h"
#include "pm_plan.h"

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pm {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

inline void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 4;
    check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
}

struct Twiddles;        // device twiddle tables shared between plans (pm_api.cpp)
struct FullTables;      // the full K x N tables, built when a plan needs them (pm_api.cpp)
struct HostPipeline;    // host transfer pipeline of a plan (pm_host.cpp)
void destroy_pipeline(HostPipeline* hp);

}  // namespace pm

// The object behind a pm_plan handle.
struct pm_plan {
    pm::HostPlan host;
    pm::DevPlan dev{};
    int form = PM_FORM_PLAIN;
    std::shared_ptr<pm::Twiddles> tw;
    std::shared_ptr<pm::FullTables> full;   // null unless table-mode rows or the generic kernels need them
    std::vector<void*> owned;            // device allocations
    uint32_t* A = nullptr;               // K x N residues of a, then y
    uint32_t* B = nullptr;               // K x N residues of b
    const pm::ProductDesc* desc = nullptr;   // its tables as batched launches read them (device)
    pm::HostPipeline* pipe = nullptr;    // host-pointer calls (created on first use)
    hipStream_t stream = nullptr;
    hipEvent_t ev[PM_NPHASES + 1] = {};    // kernel start events of the phase profile, and spares
    hipEvent_t ev_stop[PM_NPHASES] = {};   // kernel end events of the phase profile
    bool mfma_ready = false;

    ~pm_plan() {
        if (stream) (void)hipStreamSynchronize(stream);
        pm::destroy_pipeline(pipe);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto e : ev_stop)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        for (void* p : owned) (void)hipFree(p);
    }

    template <class T>
    T* upload(const std::vector<T>& v) {
        T* d = static_cast<T*>(pm::dev_alloc(v.size() * sizeof(T)));
        owned.push_back(d);
        if (!v.empty()) pm::check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
        return d;
    }

    size_t words() const { return static_cast<size_t>(host.N) * host.L; }
    size_t residues() const { return static_cast<size_t>(host.K) * host.N; }

    void ensure_mfma() {
        if (mfma_ready) return;
        pm::prepare_mfma(host, dev, [this](const void* src, size_t bytes) {
            void* d = pm::dev_alloc(bytes);
            owned.push_back(d);
            pm::check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "upload mfma tables");
            return d;
        });
        mfma_ready = true;
    }

    // The device copy of this plan's per-product tables (after ensure_mfma).
    void upload_desc() {
        pm::ProductDesc d{};
        d.K = dev.K;
        d.L = dev.L;
        d.dec_ksteps = dev.dec_ksteps;
        d.crt_ksteps = dev.crt_ksteps;
        d.dec_magic = dev.dec_magic;
        d.dec_digits = dev.dec_digits;
        d.dec_init = dev.dec_init;
        d.dec_digits_b = dev.dec_digits_b;
        d.dec_init_b = dev.dec_init_b;
        d.crt_tab = dev.crt_tab;
        d.crt_fc = dev.crt_fc;
        d.nq = dev.nq;
        d.crt_lc = dev.crt_lc;
        desc = upload(std::vector<pm::ProductDesc>{d});
    }
};


namespace pm {

// The phases of one product on device pointers, enqueued on st (pm_api.cpp):
// decompose a, decompose b, forward columns of A and of B, fused rows, inverse
// columns, CRT. run_phase enqueues phase i alone.
void run_device(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st);
void run_phase(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st, int i);
// The same with the residues of a and b in caller-chosen K x N arrays A and B.
void run_device_in(pm_plan* pl, uint32_t* A, uint32_t* B, const uint32_t* a, const uint32_t* b, uint32_t* c,
                   hipStream_t st);

// Issue-rate probe (pm_probe.hip): wall time of probe `kind` and the
// instructions each SIMD issued.
void probe_issue(int kind, int iters, double* ms, double* valu_per_simd, double* mfma_per_simd, double* clock_mhz);

// Words of the allocation-free entry point's device workspace (pm_scratch.hip):
// 72 MiB, the two K x N residue arrays of any product up to N = 65536 with a
// 2048-bit q (2 x 134 x 65536 words).
constexpr size_t kScratchWords = size_t(72) << 18;
// Its device address on the current device, looked up once per device.
uint32_t* scratch_on_current_device();

// A product on host pointers through the plan's pipeline (pm_host.cpp).
void host_multiply(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c);
// Products on host pointers, up to `depth` at once through their plans' pipelines.
void host_multiply_batch(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                         uint32_t* const* c, int depth);
// The same with the answer's memory supplied late (fherma_run): size_out runs
// on the pipeline's helper thread while the inputs stream, and get_out, called
// after it has finished, gives the memory just before the first answer word is
// written.
void host_multiply_sized(pm_plan* pl, const uint32_t* a, const uint32_t* b, std::function<void()> size_out,
                         std::function<uint32_t*()> get_out);
// Times the host modes on dummy operands and keeps the fastest; returns its
// median milliseconds per product.
double host_autotune(pm_plan* pl);
void host_set_modes(pm_plan* pl, int in_mode, int out_mode);
int host_get_mode(const pm_plan* pl);       // the input side's mode
int host_get_out_mode(const pm_plan* pl);
// The host knobs of the environment (FHERMA_IN, FHERMA_OUT, FHERMA_CHUNKS, ...,
// PM_HOST_MODE); returns whether fherma_init should run the timed choice
// (not when a mode is given or FHERMA_AUTOTUNE / PM_AUTOTUNE is 0).
bool host_configure_from_env(pm_plan* pl);

}  // namespace pm
