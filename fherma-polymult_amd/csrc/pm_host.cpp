// Host transfer pipeline for products on host memory (pm_multiply, fherma_run).
//
// Reference counterpart: solve.cu:282-706 (copy modes, CopyPool, chunk spans)
// and 817-862 (the timed mode choice at init). Three modes:
//   staged  the caller's a and b are copied chunk by chunk into pinned staging
//           (all copy threads on each chunk), sent with hipMemcpyAsync on a copy
//           stream, and each chunk is decomposed as soon as it has landed; a's
//           forward column pass runs while b is still on its way. The CRT runs
//           chunk by chunk too, each chunk's D2H starts when its CRT ends, and
//           the host copies it out as soon as it has landed.
//   mapped  the same chunks, but the decompose reads the pinned buffer through
//           its device mapping and the CRT writes the answer there: no DMA.
//   direct  hipMemcpyAsync from and to the caller's pageable memory.
// The timed choice covers the mode and, for staged and mapped, the number of
// chunks per operand (1, 2 or 4; PM_HOST_CHUNKS fixes it).
// XNACK-based access to pageable memory (the reference's hmm mode) is not used:
// it needs xnack+ code objects and HSA_XNACK=1 on the machine.
#include "../../include/fherma_polymult.h"
#include "pm_state.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace pm {
namespace {

// memcpy split into 1 MiB pieces that the calling thread and the workers claim
// from a shared ticket. The ticket carries the job's generation in its high
// half, so a worker still holding an older job can never claim a piece of the
// current one. Workers sleep between copies.
class CopyPool {
public:
    explicit CopyPool(int workers) {
        for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { loop(); });
    }

    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        wake_.notify_all();
        for (auto& t : threads_) t.join();
    }

    void copy(void* dst, const void* src, size_t bytes) {
        if (threads_.empty() || bytes < 2 * kPiece) {
            std::memcpy(dst, src, bytes);
            return;
        }
        Job job;
        {
            std::lock_guard<std::mutex> lk(m_);
            job = Job{++generation_, static_cast<char*>(dst), static_cast<const char*>(src), bytes,
                      (bytes + kPiece - 1) / kPiece};
            job_ = job;
            done_.store(0);
            ticket_.store(job.gen << 32);
        }
        wake_.notify_all();
        work(job);
        std::unique_lock<std::mutex> lk(m_);
        finished_.wait(lk, [&] { return done_.load() == job.pieces; });
    }

private:
    static constexpr size_t kPiece = size_t(1) << 20;

    struct Job {
        uint64_t gen = 0;
        char* dst = nullptr;
        const char* src = nullptr;
        size_t bytes = 0, pieces = 0;
    };

    void work(const Job& job) {
        for (;;) {
            uint64_t v = ticket_.load();
            do {
                if ((v >> 32) != job.gen || (v & 0xFFFFFFFFu) >= job.pieces) return;
            } while (!ticket_.compare_exchange_weak(v, v + 1));
            const size_t off = (v & 0xFFFFFFFFu) * kPiece;
            std::memcpy(job.dst + off, job.src + off, std::min(kPiece, job.bytes - off));
            if (done_.fetch_add(1) + 1 == job.pieces) {
                std::lock_guard<std::mutex> lk(m_);
                finished_.notify_all();
            }
        }
    }

    void loop() {
        uint64_t seen = 0;
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(m_);
                wake_.wait(lk, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                job = job_;
            }
            work(job);
        }
    }

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable wake_, finished_;
    bool stop_ = false;
    uint64_t generation_ = 0;
    Job job_;
    std::atomic<uint64_t> ticket_{0};
    std::atomic<size_t> done_{0};
};

int copy_workers() {
    if (const char* v = std::getenv("PM_COPY_THREADS")) return std::max(0, std::atoi(v) - 1);
    const int cores = static_cast<int>(std::thread::hardware_concurrency());
    return std::max(0, std::min(8, cores / 2) - 1);
}

struct Span {
    uint32_t k0, count;
};

// n near-equal coefficient ranges whose boundaries are multiples of 64 (the
// decompose and CRT workgroup tile).
std::vector<Span> spans(uint32_t N, int n) {
    std::vector<Span> out;
    const uint32_t tiles = (N + 63) / 64;
    n = std::max(1, std::min<int>(n, static_cast<int>(tiles)));
    uint32_t k0 = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t t1 = static_cast<uint32_t>((static_cast<uint64_t>(tiles) * (i + 1)) / n);
        const uint32_t k1 = std::min(N, t1 * 64);
        if (k1 > k0) out.push_back(Span{k0, k1 - k0});
        k0 = k1;
    }
    return out;
}

}  // namespace

constexpr int kMaxChunks = 16;

struct HostPipeline {
    int mode = PM_HOST_STAGED;
    int chunks = 2;             // coefficient chunks per operand (staged and mapped)
    bool chunks_fixed = false;  // PM_HOST_CHUNKS given: the timed choice leaves it alone
    uint32_t N = 0;
    std::vector<Span> in, out;
    size_t words = 0;
    uint32_t *h_a = nullptr, *h_b = nullptr, *h_c = nullptr;   // pinned, mapped
    uint32_t *m_a = nullptr, *m_b = nullptr, *m_c = nullptr;   // their device addresses
    uint32_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;   // device limbs
    hipStream_t copy = nullptr;
    std::vector<hipEvent_t> ev_a, ev_b, ev_crt, ev_out;
    std::unique_ptr<CopyPool> pool;

    ~HostPipeline() {
        for (auto* v : {&ev_a, &ev_b, &ev_crt, &ev_out})
            for (auto e : *v) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        for (auto* p : {h_a, h_b, h_c})
            if (p) (void)hipHostFree(p);
        for (auto* p : {d_a, d_b, d_c})
            if (p) (void)hipFree(p);
    }
};

void destroy_pipeline(HostPipeline* hp) { delete hp; }

static void set_chunks(HostPipeline& hp, int n) {
    hp.chunks = std::max(1, std::min(n, kMaxChunks));
    hp.in = spans(hp.N, hp.chunks);
    hp.out = spans(hp.N, hp.chunks);
}

namespace {

HostPipeline& pipeline(pm_plan* pl) {
    if (pl->pipe) return *pl->pipe;
    std::unique_ptr<HostPipeline> hp(new HostPipeline());
    hp->words = pl->words();
    const size_t bytes = hp->words * sizeof(uint32_t);
    const char* env = std::getenv("PM_HOST_CHUNKS");
    hp->N = pl->host.N;
    hp->chunks_fixed = env != nullptr;
    set_chunks(*hp, env ? std::atoi(env) : 2);
    for (auto pr : {std::make_pair(&hp->h_a, &hp->m_a), std::make_pair(&hp->h_b, &hp->m_b),
                    std::make_pair(&hp->h_c, &hp->m_c)}) {
        check(hipHostMalloc(reinterpret_cast<void**>(pr.first), bytes, hipHostMallocMapped), "hipHostMalloc");
        check(hipHostGetDevicePointer(reinterpret_cast<void**>(pr.second), *pr.first, 0), "hipHostGetDevicePointer");
    }
    for (auto** p : {&hp->d_a, &hp->d_b, &hp->d_c}) *p = static_cast<uint32_t*>(dev_alloc(bytes));
    check(hipStreamCreateWithFlags(&hp->copy, hipStreamNonBlocking), "hipStreamCreate");
    auto events = [](std::vector<hipEvent_t>& v, size_t n) {
        v.resize(n);
        for (auto& e : v) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    };
    events(hp->ev_a, kMaxChunks);
    events(hp->ev_b, kMaxChunks);
    events(hp->ev_crt, kMaxChunks);
    events(hp->ev_out, kMaxChunks);
    hp->pool.reset(new CopyPool(copy_workers()));
    pl->pipe = hp.release();
    return *pl->pipe;
}

// b goes into B scaled by the pointwise constant (b_side), as in run_device.
void decompose_range(pm_plan* pl, const uint32_t* limbs, uint32_t* res, const Span& s, hipStream_t st) {
    const DevPlan P = res == pl->B ? b_side(pl->dev) : pl->dev;
    if (pl->form == PM_FORM_MFMA)
        launch_decompose_mfma_range(P, limbs, res, s.k0, s.count, st, true);
    else
        launch_decompose_valu_range(P, limbs, res, s.k0, s.count, st);
}

void crt_range(pm_plan* pl, const uint32_t* Y, uint32_t* c, const Span& s, hipStream_t st) {
    if (pl->form == PM_FORM_MFMA)
        launch_crt_mfma_range(pl->dev, Y, c, s.k0, s.count, st);
    else
        launch_crt_valu_range(pl->dev, Y, c, s.k0, s.count, st);
}

// The device part after both operands are decomposed.
void ntt_and_product(pm_plan* pl, hipStream_t st) {
    launch_ntt_cols(pl->dev, pl->B, true, st);
    launch_ntt_rows_fused(pl->dev, pl->A, pl->B, st);
    launch_ntt_cols(pl->dev, pl->A, false, st);
}

void run_staged(pm_plan* pl, HostPipeline& hp, const uint32_t* a, const uint32_t* b, uint32_t* c, bool mapped) {
    const size_t L = static_cast<size_t>(pl->host.L);
    hipStream_t comp = pl->stream;
    auto send = [&](const uint32_t* src, uint32_t* pin, uint32_t* pin_dev, uint32_t* dev, uint32_t* res,
                    std::vector<hipEvent_t>& landed) {
        for (size_t ch = 0; ch < hp.in.size(); ++ch) {
            const Span& s = hp.in[ch];
            const size_t off = s.k0 * L, bytes = s.count * L * sizeof(uint32_t);
            hp.pool->copy(pin + off, src + off, bytes);
            if (mapped) {
                decompose_range(pl, pin_dev, res, s, comp);
                continue;
            }
            check(hipMemcpyAsync(dev + off, pin + off, bytes, hipMemcpyHostToDevice, hp.copy), "copy in");
            check(hipEventRecord(landed[ch], hp.copy), "hipEventRecord");
            check(hipStreamWaitEvent(comp, landed[ch], 0), "hipStreamWaitEvent");
            decompose_range(pl, dev, res, s, comp);
        }
    };
    send(a, hp.h_a, hp.m_a, hp.d_a, pl->A, hp.ev_a);
    launch_ntt_cols(pl->dev, pl->A, true, comp);   // a's column pass while b arrives
    send(b, hp.h_b, hp.m_b, hp.d_b, pl->B, hp.ev_b);
    ntt_and_product(pl, comp);
    for (size_t ch = 0; ch < hp.out.size(); ++ch) {
        const Span& s = hp.out[ch];
        const size_t off = s.k0 * L, bytes = s.count * L * sizeof(uint32_t);
        if (mapped) {
            crt_range(pl, pl->A, hp.m_c, s, comp);
            check(hipEventRecord(hp.ev_out[ch], comp), "hipEventRecord");
            continue;
        }
        crt_range(pl, pl->A, hp.d_c, s, comp);
        check(hipEventRecord(hp.ev_crt[ch], comp), "hipEventRecord");
        check(hipStreamWaitEvent(hp.copy, hp.ev_crt[ch], 0), "hipStreamWaitEvent");
        check(hipMemcpyAsync(hp.h_c + off, hp.d_c + off, bytes, hipMemcpyDeviceToHost, hp.copy), "copy out");
        check(hipEventRecord(hp.ev_out[ch], hp.copy), "hipEventRecord");
    }
    check(hipGetLastError(), "kernel launch");
    for (size_t ch = 0; ch < hp.out.size(); ++ch) {
        const Span& s = hp.out[ch];
        const size_t off = s.k0 * L, bytes = s.count * L * sizeof(uint32_t);
        check(hipEventSynchronize(hp.ev_out[ch]), "hipEventSynchronize");
        hp.pool->copy(c + off, hp.h_c + off, bytes);
    }
}

void run_direct(pm_plan* pl, HostPipeline& hp, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    const size_t bytes = hp.words * sizeof(uint32_t);
    hipStream_t comp = pl->stream;
    check(hipMemcpyAsync(hp.d_a, a, bytes, hipMemcpyHostToDevice, comp), "copy a");
    check(hipMemcpyAsync(hp.d_b, b, bytes, hipMemcpyHostToDevice, comp), "copy b");
    run_device(pl, hp.d_a, hp.d_b, hp.d_c, comp);
    check(hipMemcpyAsync(c, hp.d_c, bytes, hipMemcpyDeviceToHost, comp), "copy c");
    check(hipStreamSynchronize(comp), "hipStreamSynchronize");
}

}  // namespace

void host_multiply(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    HostPipeline& hp = pipeline(pl);
    switch (hp.mode) {
        case PM_HOST_MAPPED: run_staged(pl, hp, a, b, c, true); break;
        case PM_HOST_DIRECT: run_direct(pl, hp, a, b, c); break;
        default: run_staged(pl, hp, a, b, c, false); break;
    }
}

void host_set_mode(pm_plan* pl, int mode) { pipeline(pl).mode = mode; }

int host_get_mode(const pm_plan* pl) { return pl->pipe ? pl->pipe->mode : PM_HOST_STAGED; }

double host_autotune(pm_plan* pl) {
    HostPipeline& hp = pipeline(pl);
    // Dummy operands: the setup never sees a or b (timing does not depend on them).
    std::vector<uint32_t> a(hp.words), b(hp.words), c(hp.words);
    for (size_t i = 0; i < hp.words; ++i) {
        a[i] = static_cast<uint32_t>(i * 2654435761u) >> 1;
        b[i] = static_cast<uint32_t>(i * 2246822519u) >> 1;
    }
    // (mode, chunks) in timing order; the first, staged in the default chunk
    // count, stands unless another is clearly faster. The chunk count matters
    // only to staged and mapped; PM_HOST_CHUNKS fixes it.
    const int chunk0 = hp.chunks;
    std::vector<std::pair<int, int>> cands{{PM_HOST_STAGED, chunk0}};
    for (int m : {PM_HOST_STAGED, PM_HOST_MAPPED})
        for (int n : {1, 2, 4})
            if (!hp.chunks_fixed && !(m == PM_HOST_STAGED && n == chunk0)) cands.emplace_back(m, n);
    if (hp.chunks_fixed) cands.emplace_back(PM_HOST_MAPPED, chunk0);
    cands.emplace_back(PM_HOST_DIRECT, chunk0);
    double best = 1e30;
    std::pair<int, int> pick = cands.front();
    for (size_t i = 0; i < cands.size(); ++i) {
        hp.mode = cands[i].first;
        set_chunks(hp, cands[i].second);
        host_multiply(pl, a.data(), b.data(), c.data());   // first use of the setting, untimed
        std::vector<double> t;
        for (int r = 0; r < 5; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            host_multiply(pl, a.data(), b.data(), c.data());
            t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        std::sort(t.begin(), t.end());
        const double med = t[2];
        if (std::getenv("PM_HOST_TUNE_LOG"))
            std::fprintf(stderr, "host mode %d, %d chunks: %.3f ms (median of 5)\n", cands[i].first, cands[i].second, med);
        if (i == 0 ? med < best : med < 0.95 * best) {
            best = med;
            pick = cands[i];
        }
    }
    hp.mode = pick.first;
    set_chunks(hp, pick.second);
    return best;
}

}  // namespace pm
