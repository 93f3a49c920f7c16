// Host transfer pipeline for products on host memory (pm_multiply, fherma_run).
//
// Reference counterpart: solve.cu:282-706 (copy modes, CopyPool, Helper, chunk
// spans) and 817-862 (the timed mode choice at init). The input side (a and b)
// and the output side (c) each use one of three modes:
//   staged  chunk by chunk through pinned staging: the copy threads move a
//           chunk into pinned memory, hipMemcpyAsync sends it on a copy stream,
//           and it is decomposed as soon as it has landed (a's forward column
//           pass runs while b is still on its way); on the way out each CRT
//           chunk's D2H starts when that chunk is done and the host copies it
//           out as soon as it has landed.
//   mapped  the same chunks, but the decompose reads the pinned buffer through
//           its device mapping, and the CRT writes there: no DMA.
//   direct  hipMemcpyAsync from and to the caller's pageable memory.
// fherma_run's answer is sized (and zero-filled, as std::vector does) on a
// helper thread while the inputs stream, as the reference does
// (solve.cu:907-918); the first output chunk waits for it.
// XNACK-based access to pageable memory (the reference's hmm and stream modes)
// needs xnack+ code objects and HSA_XNACK=1, which this build does not assume;
// like the reference on a device without pageable access, those names fall
// back to staged.
#include "../../include/fherma_polymult.h"
#include "pm_state.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <string>
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

// One persistent thread that runs one job at a time for the caller: it sizes
// fherma_run's answer while the inputs stream (reference solve.cu:419-481,
// 907-918). wait() returns when the job is done and rethrows its exception.
class Helper {
public:
    Helper() : thread_([this] { loop(); }) {}
    ~Helper() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        thread_.join();
    }

    void run(std::function<void()> job) {
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = std::move(job);
            pending_ = true;
            error_ = nullptr;
        }
        cv_.notify_all();
    }

    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return !pending_; });
        if (error_) {
            std::exception_ptr e = error_;
            error_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || pending_; });
            if (stop_) return;
            std::function<void()> job = std::move(job_);
            lk.unlock();
            std::exception_ptr err;
            try {
                job();
            } catch (...) {
                err = std::current_exception();
            }
            lk.lock();
            error_ = err;
            pending_ = false;
            done_.notify_all();
        }
    }

    std::mutex m_;
    std::condition_variable cv_, done_;
    std::function<void()> job_;
    bool pending_ = false, stop_ = false;
    std::exception_ptr error_;
    std::thread thread_;
};

const char* env2(const char* a, const char* b) {
    const char* v = std::getenv(a);
    return v ? v : std::getenv(b);
}

int copy_workers() {
    if (const char* v = env2("FHERMA_THREADS", "PM_COPY_THREADS")) return std::max(0, std::atoi(v) - 1);
    const int cores = static_cast<int>(std::thread::hardware_concurrency());
    return std::max(0, std::min(8, cores / 2) - 1);
}

struct Span {
    uint32_t k0, count;
};

// Coefficient ranges in proportion to the weights w, every boundary but the
// last on a multiple of 64 (the decompose and CRT tile); empty ranges dropped.
std::vector<Span> spans(uint32_t N, const std::vector<double>& w) {
    double total = 0;
    for (double x : w) total += x;
    std::vector<Span> out;
    double acc = 0;
    uint32_t k0 = 0;
    for (size_t i = 0; i < w.size(); ++i) {
        acc += w[i];
        uint32_t k1 = i + 1 == w.size() ? N : static_cast<uint32_t>(std::llround(acc / total * N / 64.0)) * 64u;
        k1 = std::min(std::max(k1, k0), N);
        if (k1 > k0) out.push_back(Span{k0, k1 - k0});
        k0 = k1;
    }
    return out;
}

std::vector<double> equal(int n) { return std::vector<double>(static_cast<size_t>(std::max(1, n)), 1.0); }

// FHERMA_A_CHUNKS style weights "w,w,..", positive (reference solve.cu:527-540).
bool env_weights(const char* var, std::vector<double>& w) {
    const char* v = std::getenv(var);
    if (!v) return false;
    w.clear();
    for (const char* p = v; *p;) {
        char* end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p || !(x > 0)) throw std::runtime_error(std::string(var) + ": positive weights, comma separated");
        w.push_back(x);
        p = *end == ',' ? end + 1 : end;
    }
    if (w.empty()) throw std::runtime_error(std::string(var) + ": positive weights, comma separated");
    return true;
}

}  // namespace

constexpr int kMaxChunks = 16;

struct HostPipeline {
    int in_mode = PM_HOST_STAGED, out_mode = PM_HOST_STAGED;
    int chunks = 2;              // coefficient chunks per operand, both sides (when no weights are given)
    bool chunks_fixed = false;   // set by the environment: the timed choice leaves the layout alone
    bool profile = false;        // FHERMA_PROFILE: per-phase times on stderr, every call
    bool tune_log = false;
    uint32_t N = 0;
    std::vector<Span> a_sp, b_sp, c_sp;
    size_t words = 0;
    uint32_t *h_a = nullptr, *h_b = nullptr, *h_c = nullptr;   // pinned, mapped
    uint32_t *m_a = nullptr, *m_b = nullptr, *m_c = nullptr;   // their device addresses
    uint32_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;   // device limbs
    hipStream_t copy = nullptr;
    std::vector<hipEvent_t> ev_a, ev_b, ev_crt, ev_out;
    hipEvent_t prof[5] = {};
    std::unique_ptr<CopyPool> pool;
    std::unique_ptr<Helper> helper;

    ~HostPipeline() {
        helper.reset();
        pool.reset();
        for (auto* v : {&ev_a, &ev_b, &ev_crt, &ev_out})
            for (auto e : *v) (void)hipEventDestroy(e);
        for (auto e : prof)
            if (e) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
        for (auto* p : {h_a, h_b, h_c})
            if (p) (void)hipHostFree(p);
        for (auto* p : {d_a, d_b, d_c})
            if (p) (void)hipFree(p);
    }
};

void destroy_pipeline(HostPipeline* hp) { delete hp; }

namespace {

void set_chunks(HostPipeline& hp, int n) {
    hp.chunks = std::max(1, std::min(n, kMaxChunks));
    hp.a_sp = hp.b_sp = hp.c_sp = spans(hp.N, equal(hp.chunks));
}

// The chunk layout from the environment (reference solve.cu:686-700):
// FHERMA_CHUNKS (or PM_HOST_CHUNKS) input chunks, FHERMA_OUT_CHUNKS output
// chunks, FHERMA_A/B/C_CHUNKS relative weights per chunk. Any of them fixes
// the layout for the timed choice.
void chunks_from_env(HostPipeline& hp) {
    const char* in = env2("FHERMA_CHUNKS", "PM_HOST_CHUNKS");
    const char* out = std::getenv("FHERMA_OUT_CHUNKS");
    const int n_in = in ? std::max(1, std::atoi(in)) : hp.chunks;
    const int n_out = out ? std::max(1, std::atoi(out)) : n_in;
    std::vector<double> wa = equal(n_in), wb = equal(n_in), wc = equal(n_out);
    const bool any = env_weights("FHERMA_A_CHUNKS", wa) | env_weights("FHERMA_B_CHUNKS", wb) |
                     env_weights("FHERMA_C_CHUNKS", wc);
    for (const auto* w : {&wa, &wb, &wc})
        if (w->size() > static_cast<size_t>(kMaxChunks))
            throw std::runtime_error("at most " + std::to_string(kMaxChunks) + " chunks per operand");
    hp.chunks_fixed = in || out || any;
    if (!hp.chunks_fixed) return;
    hp.chunks = std::min(n_in, kMaxChunks);
    hp.a_sp = spans(hp.N, wa);
    hp.b_sp = spans(hp.N, wb);
    hp.c_sp = spans(hp.N, wc);
}

HostPipeline& pipeline(pm_plan* pl) {
    if (pl->pipe) return *pl->pipe;
    std::unique_ptr<HostPipeline> hp(new HostPipeline());
    hp->words = pl->words();
    const size_t bytes = hp->words * sizeof(uint32_t);
    hp->N = pl->host.N;
    set_chunks(*hp, 2);
    chunks_from_env(*hp);
    hp->profile = std::getenv("FHERMA_PROFILE") != nullptr;
    hp->tune_log = env2("FHERMA_AUTOTUNE_LOG", "PM_HOST_TUNE_LOG") != nullptr;
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
    for (auto& e : hp->prof) check(hipEventCreate(&e), "hipEventCreate");
    hp->pool.reset(new CopyPool(copy_workers()));
    hp->helper.reset(new Helper());
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

// The answer's memory, supplied late: sized on the helper, fetched when the
// first output word is about to be written.
struct Output {
    HostPipeline& hp;
    std::function<uint32_t*()> get;
    uint32_t* c = nullptr;
    bool sizing = false;
    uint32_t* ptr() {
        if (!c) {
            if (sizing) {
                sizing = false;
                hp.helper->wait();
            }
            c = get();
        }
        return c;
    }
    ~Output() {
        if (sizing) {   // the job writes into the caller's answer: wait however the call ends
            try {
                hp.helper->wait();
            } catch (...) {
            }
        }
    }
};

void send_operand(pm_plan* pl, HostPipeline& hp, const uint32_t* src, uint32_t* pin, uint32_t* pin_dev,
                  uint32_t* dev, uint32_t* res, const std::vector<Span>& sp, std::vector<hipEvent_t>& landed) {
    const size_t L = static_cast<size_t>(pl->host.L);
    hipStream_t comp = pl->stream;
    if (hp.in_mode == PM_HOST_DIRECT) {
        check(hipMemcpyAsync(dev, src, hp.words * sizeof(uint32_t), hipMemcpyHostToDevice, comp), "copy in");
        decompose_range(pl, dev, res, Span{0, hp.N}, comp);
        return;
    }
    for (size_t ch = 0; ch < sp.size(); ++ch) {
        const Span& s = sp[ch];
        const size_t off = s.k0 * L, bytes = s.count * L * sizeof(uint32_t);
        hp.pool->copy(pin + off, src + off, bytes);
        if (hp.in_mode == PM_HOST_MAPPED) {
            decompose_range(pl, pin_dev, res, s, comp);
            continue;
        }
        check(hipMemcpyAsync(dev + off, pin + off, bytes, hipMemcpyHostToDevice, hp.copy), "copy in");
        check(hipEventRecord(landed[ch], hp.copy), "hipEventRecord");
        check(hipStreamWaitEvent(comp, landed[ch], 0), "hipStreamWaitEvent");
        decompose_range(pl, dev, res, s, comp);
    }
}

void run(pm_plan* pl, HostPipeline& hp, const uint32_t* a, const uint32_t* b, Output& out) {
    const size_t L = static_cast<size_t>(pl->host.L);
    hipStream_t comp = pl->stream;
    const auto t0 = std::chrono::steady_clock::now();
    if (hp.profile) check(hipEventRecord(hp.prof[0], comp), "hipEventRecord");
    send_operand(pl, hp, a, hp.h_a, hp.m_a, hp.d_a, pl->A, hp.a_sp, hp.ev_a);
    launch_ntt_cols(pl->dev, pl->A, true, comp);   // a's column pass while b arrives
    send_operand(pl, hp, b, hp.h_b, hp.m_b, hp.d_b, pl->B, hp.b_sp, hp.ev_b);
    if (hp.profile) check(hipEventRecord(hp.prof[1], comp), "hipEventRecord");
    launch_ntt_cols(pl->dev, pl->B, true, comp);
    launch_ntt_rows_fused(pl->dev, pl->A, pl->B, comp);
    launch_ntt_cols(pl->dev, pl->A, false, comp);
    if (hp.profile) check(hipEventRecord(hp.prof[2], comp), "hipEventRecord");
    const auto t1 = std::chrono::steady_clock::now();
    if (hp.out_mode == PM_HOST_DIRECT) {
        crt_range(pl, pl->A, hp.d_c, Span{0, hp.N}, comp);
        if (hp.profile) check(hipEventRecord(hp.prof[3], comp), "hipEventRecord");
        check(hipGetLastError(), "kernel launch");
        uint32_t* c = out.ptr();
        check(hipMemcpyAsync(c, hp.d_c, hp.words * sizeof(uint32_t), hipMemcpyDeviceToHost, comp), "copy out");
        if (hp.profile) check(hipEventRecord(hp.prof[4], comp), "hipEventRecord");
        check(hipStreamSynchronize(comp), "hipStreamSynchronize");
    } else {
        const bool mapped = hp.out_mode == PM_HOST_MAPPED;
        for (size_t ch = 0; ch < hp.c_sp.size(); ++ch) {
            const Span& s = hp.c_sp[ch];
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
        if (hp.profile) check(hipEventRecord(hp.prof[3], comp), "hipEventRecord");
        check(hipGetLastError(), "kernel launch");
        for (size_t ch = 0; ch < hp.c_sp.size(); ++ch) {
            const Span& s = hp.c_sp[ch];
            const size_t off = s.k0 * L, bytes = s.count * L * sizeof(uint32_t);
            check(hipEventSynchronize(hp.ev_out[ch]), "hipEventSynchronize");
            hp.pool->copy(out.ptr() + off, hp.h_c + off, bytes);
        }
        if (hp.profile) {
            check(hipEventRecord(hp.prof[4], mapped ? comp : hp.copy), "hipEventRecord");
            check(hipEventSynchronize(hp.prof[4]), "hipEventSynchronize");
        }
    }
    if (hp.profile) {
        // GPU phases from the events, then the host's own timeline (reference solve.cu:650-659)
        float ms[4] = {};
        for (int i = 0; i < 4; ++i) check(hipEventElapsedTime(&ms[i], hp.prof[i], hp.prof[i + 1]), "hipEventElapsedTime");
        const auto t2 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "gpu us: in+decompose %.1f  ntt %.1f  crt %.1f  out-tail %.1f | "
                     "host us: enqueue %.1f  drain %.1f\n", ms[0] * 1e3, ms[1] * 1e3, ms[2] * 1e3, ms[3] * 1e3,
                     std::chrono::duration<double, std::micro>(t1 - t0).count(),
                     std::chrono::duration<double, std::micro>(t2 - t1).count());
    }
}

}  // namespace

void host_multiply(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    HostPipeline& hp = pipeline(pl);
    Output out{hp, [c] { return c; }};
    run(pl, hp, a, b, out);
}

void host_multiply_sized(pm_plan* pl, const uint32_t* a, const uint32_t* b, std::function<void()> size_out,
                         std::function<uint32_t*()> get_out) {
    HostPipeline& hp = pipeline(pl);
    Output out{hp, std::move(get_out)};
    hp.helper->run(std::move(size_out));
    out.sizing = true;
    run(pl, hp, a, b, out);
}

// Products on host memory, up to `depth` at once, each through its own plan's
// pipeline (its own pinned buffers, streams and copy threads), so that one
// product's input copies overlap another's kernels and a third's output. The
// products of one plan run one after another.
void host_multiply_batch(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                         uint32_t* const* c, int depth) {
    if (count <= 0) return;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    for (int i = 0; i < count; ++i) pipeline(plans[i]);   // created here, on the caller's thread
    std::vector<std::unique_ptr<std::mutex>> locks;
    std::vector<int> lock_of(count);
    for (int i = 0; i < count; ++i) {
        lock_of[i] = i;
        for (int k = 0; k < i; ++k)
            if (plans[k] == plans[i]) {
                lock_of[i] = lock_of[k];
                break;
            }
        locks.emplace_back(new std::mutex());
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr first;
    std::mutex first_mutex;
    auto work = [&](bool own_thread) {
        try {
            if (own_thread) check(hipSetDevice(dev), "hipSetDevice");
            for (int i; !failed.load() && (i = next.fetch_add(1)) < count;) {
                std::lock_guard<std::mutex> lock(*locks[lock_of[i]]);
                host_multiply(plans[i], a[i], b[i], c[i]);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(first_mutex);
            if (!first) first = std::current_exception();
            failed = true;
        }
    };
    std::vector<std::thread> threads;
    const int n = std::max(1, std::min(depth, count));
    for (int t = 1; t < n; ++t) threads.emplace_back(work, true);
    work(false);
    for (auto& t : threads) t.join();
    if (first) std::rethrow_exception(first);
}

void host_set_modes(pm_plan* pl, int in_mode, int out_mode) {
    HostPipeline& hp = pipeline(pl);
    hp.in_mode = in_mode;
    hp.out_mode = out_mode;
}

int host_get_mode(const pm_plan* pl) { return pl->pipe ? pl->pipe->in_mode : PM_HOST_STAGED; }
int host_get_out_mode(const pm_plan* pl) { return pl->pipe ? pl->pipe->out_mode : PM_HOST_STAGED; }

namespace {

// FHERMA_IN / FHERMA_OUT names (reference solve.cu:287-296); hmm and stream
// need pageable access from the device, which this build does not assume, and
// fall back to staged as the reference does on such a device (solve.cu:729-735).
int mode_of(const char* var, const char* v) {
    if (!std::strcmp(v, "staged") || !std::strcmp(v, "hmm") || !std::strcmp(v, "stream")) return PM_HOST_STAGED;
    if (!std::strcmp(v, "mapped")) return PM_HOST_MAPPED;
    if (!std::strcmp(v, "register")) return PM_HOST_DIRECT;
    throw std::runtime_error(std::string(var) + " must be staged, register, mapped, hmm or stream");
}

}  // namespace

bool host_configure_from_env(pm_plan* pl) {
    HostPipeline& hp = pipeline(pl);
    const char* fin = std::getenv("FHERMA_IN");
    const char* fout = std::getenv("FHERMA_OUT");
    const char* pm = std::getenv("PM_HOST_MODE");
    if (fin || fout) {
        hp.in_mode = fin ? mode_of("FHERMA_IN", fin) : PM_HOST_STAGED;
        hp.out_mode = fout ? mode_of("FHERMA_OUT", fout) : PM_HOST_STAGED;
        return false;
    }
    if (pm) {
        const int m = !std::strcmp(pm, "staged")   ? PM_HOST_STAGED
                      : !std::strcmp(pm, "mapped") ? PM_HOST_MAPPED
                      : !std::strcmp(pm, "direct") ? PM_HOST_DIRECT
                                                   : -1;
        if (m < 0) throw std::runtime_error("PM_HOST_MODE must be staged, mapped or direct");
        hp.in_mode = hp.out_mode = m;
        return false;
    }
    const char* at = env2("FHERMA_AUTOTUNE", "PM_AUTOTUNE");
    return !(at && !std::strcmp(at, "0"));
}

double host_autotune(pm_plan* pl) {
    HostPipeline& hp = pipeline(pl);
    // Dummy operands: the setup never sees a or b (timing does not depend on them).
    std::vector<uint32_t> a(hp.words), b(hp.words);
    for (size_t i = 0; i < hp.words; ++i) {
        a[i] = static_cast<uint32_t>(i * 2654435761u) >> 1;
        b[i] = static_cast<uint32_t>(i * 2246822519u) >> 1;
    }
    // (in mode, out mode, chunks) in timing order; the first, staged both ways
    // in the default chunk count, stands unless another is clearly faster.
    // The chunk count matters only to staged and mapped; an environment layout
    // fixes it.
    struct Cand {
        int in, out, chunks;
    };
    const int chunk0 = hp.chunks;
    std::vector<Cand> cands{{PM_HOST_STAGED, PM_HOST_STAGED, chunk0}};
    for (int m : {PM_HOST_STAGED, PM_HOST_MAPPED})
        for (int n : {1, 2, 4})
            if (!hp.chunks_fixed && !(m == PM_HOST_STAGED && n == chunk0)) cands.push_back({m, m, n});
    if (hp.chunks_fixed) cands.push_back({PM_HOST_MAPPED, PM_HOST_MAPPED, chunk0});
    cands.push_back({PM_HOST_MAPPED, PM_HOST_STAGED, chunk0});
    cands.push_back({PM_HOST_STAGED, PM_HOST_MAPPED, chunk0});
    cands.push_back({PM_HOST_DIRECT, PM_HOST_DIRECT, chunk0});
    // Each timed call does what fherma_run does (the answer sized on the helper
    // while the inputs stream), after a write over 4 x the operand size so that
    // no run finds the previous one's data in the host caches (reference
    // solve.cu:830-841); median of 7.
    std::vector<char> flush(4 * hp.words * sizeof(uint32_t));
    struct Quiet {   // the phase times are for the caller's calls only
        HostPipeline& hp;
        bool was;
        ~Quiet() { hp.profile = was; }
    } quiet{hp, hp.profile};
    hp.profile = false;
    double best = 1e30;
    Cand pick = cands.front();
    for (size_t i = 0; i < cands.size(); ++i) {
        hp.in_mode = cands[i].in;
        hp.out_mode = cands[i].out;
        if (!hp.chunks_fixed) set_chunks(hp, cands[i].chunks);
        std::vector<double> t;
        for (int r = -1; r < 7; ++r) {   // r = -1: first use of the setting, untimed
            std::fill(flush.begin(), flush.end(), static_cast<char>(r));
            std::vector<uint32_t> c;
            const auto t0 = std::chrono::steady_clock::now();
            host_multiply_sized(pl, a.data(), b.data(), [&c, &hp] { c.resize(hp.words); }, [&c] { return c.data(); });
            if (r >= 0)
                t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        std::sort(t.begin(), t.end());
        const double med = t[3];
        if (hp.tune_log)
            std::fprintf(stderr, "host mode in %d out %d, %d chunks: %.3f ms (median of 7)\n", cands[i].in,
                         cands[i].out, cands[i].chunks, med);
        if (i == 0 ? med < best : med < 0.95 * best) {
            best = med;
            pick = cands[i];
        }
    }
    hp.in_mode = pick.in;
    hp.out_mode = pick.out;
    if (!hp.chunks_fixed) set_chunks(hp, pick.chunks);
    return best;
}

}  // namespace pm
