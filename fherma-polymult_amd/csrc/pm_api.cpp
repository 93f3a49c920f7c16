// The library's boundary: the extern "C" pm_* API (include/fherma_polymult.h)
// and the FHERMA C++ entry points fherma_init / fherma_run / fherma_free
// (include/fherma.h), over one plan per point.
//
// A plan owns: the host plan (primes, constants), the device copies of its
// tables, a K x N residue workspace per operand, N x L device limb buffers and
// pinned host staging for host-pointer calls, and one HIP stream at the default
// priority. Twiddle tables depend only on (N, primes), so plans that share
// them (every product of a batch at one N) share one device copy.
#include "../../include/fherma.h"
#include "../../include/fherma_polymult.h"
#include "pm_arith.h"
#include "pm_device.h"
#include "pm_plan.h"
#include "pm_mfma_host.h"
#include "pm_state.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;
using pm::check;
using pm::dev_alloc;
using pm::HipError;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return PM_OK;
    } catch (const HipError& e) {
        g_err = e.what();
        return PM_EHIP;
    } catch (const std::bad_alloc& e) {
        g_err = "out of memory";
        return PM_ENOMEM;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return PM_EINVAL;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        return PM_EINVAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PM_EINTERNAL;
    }
}

}  // namespace

namespace pm {

// The full K x N twiddle tables, forward and inverse (psi^bitrev(i) pairs):
// only the row pass reading its twiddles from tables and the generic NTT
// kernels (geometries without register kernels) need them, so they exist only
// while such a plan or batch holds them.
struct FullTables {
    uint2* tw = nullptr;
    uint2* itw = nullptr;
    ~FullTables() {
        if (tw) (void)hipFree(tw);
        if (itw) (void)hipFree(itw);
    }
};

// Device twiddle tables for (N, prime list), shared between plans.
struct Twiddles {
    uint2* ctw = nullptr;    // K x R: the first R pairs of each prime's forward table (column pass)
    uint2* citw = nullptr;   // inverse
    uint2* rz = nullptr;
    uint2* rzi = nullptr;
    uint32_t* rom = nullptr;
    uint32_t* romi = nullptr;
    uint8_t* colm_fwd = nullptr;   // matrix-core column pass fragments (R = 256 only)
    uint8_t* colm_inv = nullptr;
    std::mutex full_mutex;
    std::weak_ptr<FullTables> full;
    ~Twiddles() {
        for (void* p : {static_cast<void*>(ctw), static_cast<void*>(citw), static_cast<void*>(rz),
                        static_cast<void*>(rzi), static_cast<void*>(rom), static_cast<void*>(romi),
                        static_cast<void*>(colm_fwd), static_cast<void*>(colm_inv)})
            if (p) (void)hipFree(p);
    }
};

// The full tables of t's primes: the existing copy while anything holds one,
// else built and uploaded now.
std::shared_ptr<FullTables> full_tables(Twiddles& t, const HostPlan& pl) {
    std::lock_guard<std::mutex> lock(t.full_mutex);
    if (auto sp = t.full.lock()) return sp;
    auto f = std::make_shared<FullTables>();
    const size_t words = 2 * static_cast<size_t>(pl.K) * pl.N;
    std::vector<uint32_t> host(words);
    f->tw = static_cast<uint2*>(dev_alloc(words * sizeof(uint32_t)));
    f->itw = static_cast<uint2*>(dev_alloc(words * sizeof(uint32_t)));
    build_twiddles(pl, true, 0, pl.K, host.data(), pl.N);
    check(hipMemcpy(f->tw, host.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice), "upload twiddles");
    build_twiddles(pl, false, 0, pl.K, host.data(), pl.N);
    check(hipMemcpy(f->itw, host.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice), "upload twiddles");
    t.full = f;
    return f;
}

}  // namespace pm

namespace {

using pm::Twiddles;

std::mutex g_tw_mutex;
std::map<std::vector<uint32_t>, std::weak_ptr<Twiddles>> g_tw_cache;

std::shared_ptr<Twiddles> twiddles_for(const pm::HostPlan& pl) {
    std::vector<uint32_t> key(pl.p);
    key.push_back(pl.N);
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto it = g_tw_cache.find(key);
    if (it != g_tw_cache.end()) {
        if (auto sp = it->second.lock()) return sp;
    }
    auto t = std::make_shared<Twiddles>();
    const uint32_t R = 1u << pl.logR;
    const size_t words = 2 * static_cast<size_t>(pl.K) * R;
    std::vector<uint32_t> host(words);
    for (int fwd = 1; fwd >= 0; --fwd) {
        pm::build_twiddles(pl, fwd != 0, 0, pl.K, host.data(), R);
        uint2*& dst = fwd ? t->ctw : t->citw;
        dst = static_cast<uint2*>(dev_alloc(words * sizeof(uint32_t)));
        check(hipMemcpy(dst, host.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice), "upload twiddles");
    }
    const size_t zw = 2 * static_cast<size_t>(pl.K) * (size_t(1) << pl.logR) * pl.logS;
    const size_t ow = static_cast<size_t>(pl.K) * ((size_t(1) << pl.logS) / 2);
    std::vector<uint32_t> z(zw), o(ow);
    for (int fwd = 1; fwd >= 0; --fwd) {
        pm::build_row_twiddles(pl, fwd != 0, 0, pl.K, z.data(), o.data());
        // each buffer is owned by t as soon as it exists, so a failed allocation frees the others
        uint2*& dz = fwd ? t->rz : t->rzi;
        dz = static_cast<uint2*>(dev_alloc(zw * sizeof(uint32_t)));
        uint32_t*& dom = fwd ? t->rom : t->romi;
        dom = static_cast<uint32_t*>(dev_alloc(ow * sizeof(uint32_t)));
        check(hipMemcpy(dz, z.data(), zw * sizeof(uint32_t), hipMemcpyHostToDevice), "upload row twiddles");
        check(hipMemcpy(dom, o.data(), ow * sizeof(uint32_t), hipMemcpyHostToDevice), "upload row twiddles");
    }
    if (pm::col_mfma_ok(pl)) {
        std::vector<uint8_t> m;
        for (int fwd = 1; fwd >= 0; --fwd) {
            pm::build_col_mfma_tables(pl, fwd != 0, m);
            uint8_t*& dst = fwd ? t->colm_fwd : t->colm_inv;
            dst = static_cast<uint8_t*>(dev_alloc(m.size()));
            check(hipMemcpy(dst, m.data(), m.size(), hipMemcpyHostToDevice), "upload column-pass matrices");
        }
    }
    // entries whose tables every plan has released are dropped here
    for (auto e = g_tw_cache.begin(); e != g_tw_cache.end();) e = e->second.expired() ? g_tw_cache.erase(e) : std::next(e);
    g_tw_cache[key] = t;
    return t;
}

}  // namespace

namespace {

hipStream_t stream_of(pm_plan* pl, void* s) { return s ? static_cast<hipStream_t>(s) : pl->stream; }

void need(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}

// Row-pass twiddle source, the reference's rule (rns_plan.h:269-280):
// generated when the full tables (16 K N bytes) exceed 64 MB, else read from
// them; FHERMA_TWIDDLES=table|gen|auto overrides, anything else is refused.
int row_twiddle_policy(const pm::HostPlan& h) {
    if (const char* v = std::getenv("FHERMA_TWIDDLES")) {
        if (!std::strcmp(v, "gen")) return 1;
        if (!std::strcmp(v, "table")) return 0;
        need(!std::strcmp(v, "auto"), "FHERMA_TWIDDLES must be table, gen or auto");
    }
    return 16.0 * h.K * h.N > 64e6 ? 1 : 0;
}

// The column pass a new plan uses: the VALU one unless PM_NTT_FORM=mfma asks
// for the matrix-core one where its tables exist (PM_NTT_FORM=valu or auto: the
// default, which measured faster alone and equal inside the product batch;
// anything else is refused).
int ntt_form_default(const pm::DevPlan& d) {
    bool mfma = false;
    if (const char* v = std::getenv("PM_NTT_FORM")) {
        need(!std::strcmp(v, "valu") || !std::strcmp(v, "mfma") || !std::strcmp(v, "auto"),
             "PM_NTT_FORM must be valu, mfma or auto");
        mfma = std::strcmp(v, "mfma") == 0;
    }
    return mfma && d.colm_fwd && d.colm_inv ? 1 : 0;
}

// Whether a new plan's product path runs the first forward column stages
// inside the decompose and the next ones in the row pass (DevPlan::fused_cols)
// where the geometry allows: PM_FUSED_COLS=1 or auto (the default) yes, 0 no;
// anything else is refused.
int fused_cols_default(const pm::DevPlan& d) {
    bool on = true;
    if (const char* v = std::getenv("PM_FUSED_COLS")) {
        need(!std::strcmp(v, "0") || !std::strcmp(v, "1") || !std::strcmp(v, "auto"),
             "PM_FUSED_COLS must be 0, 1 or auto");
        on = std::strcmp(v, "0") != 0;
    }
    return on && pm::fused_cols_ok(d) ? 1 : 0;
}

// The host plan of a point with the reference's prime rule; `kernels`: every
// NTT pass of the geometry has a register kernel, whose faster butterfly form
// the plan then uses.
pm::HostPlan plan_for(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, bool& kernels) {
    pm::HostPlan h = pm::make_plan(N, W, q, static_cast<int>(L));
    pm::DevPlan g{};
    g.logS = h.logS;
    g.logR = h.logR;
    kernels = pm::batch_kernels_ok(g);
    return h;
}

pm_plan* make_plan_object(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q) {
    need(q != nullptr, "q is null");
    std::unique_ptr<pm_plan> pl(new pm_plan());
    bool kernels = false;
    pl->host = plan_for(N, W, L, q, kernels);
    const pm::HostPlan& h = pl->host;
    const int K = h.K;
    const int row_gen = row_twiddle_policy(h);   // before any device work

    std::vector<pm::PrimeConst> pc(K);
    for (int j = 0; j < K; ++j) {
        const uint32_t p = h.p[j];
        pc[j].p = p;
        pc[j].pneg = h.pneg[j];
        pc[j].c32 = static_cast<uint32_t>((1ull << 32) % p);
        pc[j].c32s = pm::shoup_companion(pc[j].c32, p);
        pc[j].c64 = static_cast<uint32_t>(static_cast<uint64_t>(pc[j].c32) * pc[j].c32 % p);
        pc[j].one_s = pm::shoup_companion(1, p);
        pc[j].scale = h.scale[j];
        pc[j].scale_s = h.scale_s[j];
    }
    pl->tw = twiddles_for(h);

    pm::DevPlan& d = pl->dev;
    d.N = h.N;
    d.logN = h.logN;
    d.logS = h.logS;
    d.logR = h.logR;
    d.L = h.L;
    d.K = K;
    d.row_gen = row_gen;
    d.lazy = kernels ? 1 : 0;
    d.pc = pl->upload(pc);
    d.pow2 = pl->upload(h.pow2);
    d.pow2b = pl->upload(h.pow2b);
    d.ctw = pl->tw->ctw;
    d.citw = pl->tw->citw;
    if (!kernels || row_gen == 0) {
        pl->full = pm::full_tables(*pl->tw, h);
        d.tw = pl->full->tw;
        d.itw = pl->full->itw;
    }
    d.rz = pl->tw->rz;
    d.rzi = pl->tw->rzi;
    d.rom = pl->tw->rom;
    d.romi = pl->tw->romi;
    d.colm_fwd = pl->tw->colm_fwd;
    d.colm_inv = pl->tw->colm_inv;
    d.col_mfma = ntt_form_default(d);
    d.p_recip = pl->upload(h.p_recip);
    d.M = pl->upload(h.M);
    d.negvP = pl->upload(h.negvP);
    d.q = pl->upload(h.q);
    d.mq = pl->upload(h.mq);
    d.nq = pl->upload(h.nq);

    pl->A = static_cast<uint32_t*>(dev_alloc(pl->residues() * sizeof(uint32_t)));
    pl->owned.push_back(pl->A);
    pl->B = static_cast<uint32_t*>(dev_alloc(pl->residues() * sizeof(uint32_t)));
    pl->owned.push_back(pl->B);
    check(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : pl->ev) check(hipEventCreate(&e), "hipEventCreate");
    for (auto& e : pl->ev_stop) check(hipEventCreate(&e), "hipEventCreate");
    pl->ensure_mfma();
    pl->upload_desc();
    pl->form = PM_FORM_MFMA;
    d.fused_cols = fused_cols_default(d);
    (void)pm::scratch_on_current_device();   // set-up of the allocation-free entry point on this device
    return pl.release();
}

// Whether the products of two plans can share batched launches: both in the
// matrix-core form, at one N with register NTT kernels, and the shorter prime
// list a prefix of the longer (always so for plans made by the prime rule at
// one N: it walks one sequence of primes per N and stops at a count set by q's
// width), so that one plan's per-prime constants and twiddles serve both.
bool groupable(const pm_plan* x, const pm_plan* y) {
    if (x->form != PM_FORM_MFMA || y->form != PM_FORM_MFMA) return false;
    if (x->host.N != y->host.N || !pm::batch_kernels_ok(x->dev)) return false;
    const bool xs = x->host.p.size() <= y->host.p.size();
    const std::vector<uint32_t>& shorter = xs ? x->host.p : y->host.p;
    const std::vector<uint32_t>& longer = xs ? y->host.p : x->host.p;
    return std::equal(shorter.begin(), shorter.end(), longer.begin());
}

// The column-pass form and row-pass twiddle source of a batched run, when they
// differ from those of its longest plan (-1: the plan's own). Table twiddles
// come from `full`, the full tables of that plan's primes.
struct NttChoice {
    int col_mfma = -1;
    int fused_cols = -1;
    int row_gen = -1;
    const pm::FullTables* full = nullptr;
};

// Entry k of ba as entry m of sub.
void copy_entry(const pm::BatchArgs& ba, int k, pm::BatchArgs& sub, int m) {
    sub.a[m] = ba.a[k];
    sub.b[m] = ba.b[k];
    sub.c[m] = ba.c[k];
    sub.A[m] = ba.A[k];
    sub.B[m] = ba.B[k];
    sub.d[m] = ba.d[k];
    sub.K[m] = ba.K[k];
}

// Launches go(sub, m, key) once per distinct key among the n entries of ba, sub
// holding the m entries with that key.
template <class Go>
void by_key(const pm::BatchArgs& ba, int n, const std::vector<int>& key, Go go) {
    std::vector<char> done(n, 0);
    for (int k = 0; k < n; ++k) {
        if (done[k]) continue;
        pm::BatchArgs sub{};
        int m = 0;
        for (int t = k; t < n; ++t)
            if (!done[t] && key[t] == key[k]) {
                copy_entry(ba, t, sub, m++);
                done[t] = 1;
            }
        go(sub, m, key[k]);
    }
}

// The products idx (pairwise groupable, distinct plans) as batched launches on
// st, kBatchMax products per launch, most primes first: decompose of every a
// and b (one launch per decompose K-step count), forward columns of every A and
// B, fused rows, inverse columns, CRT (one launch per CRT K-step count). The
// plan with the most primes of each launch supplies the geometry, per-prime
// constants and twiddles.
void run_batched(const std::vector<int>& idx, pm_plan* const* plans, const uint32_t* const* a,
                 const uint32_t* const* b, uint32_t* const* c, hipStream_t st, const NttChoice& ch = {}) {
    std::vector<int> order(idx);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return plans[x]->host.K > plans[y]->host.K; });
    for (size_t i0 = 0; i0 < order.size(); i0 += pm::kBatchMax) {
        const int n = static_cast<int>(std::min(order.size() - i0, static_cast<size_t>(pm::kBatchMax)));
        pm::DevPlan P = plans[order[i0]]->dev;
        if (ch.col_mfma >= 0) P.col_mfma = ch.col_mfma;
        if (ch.row_gen >= 0) P.row_gen = ch.row_gen;
        if (ch.fused_cols >= 0) P.fused_cols = ch.fused_cols && pm::fused_cols_ok(P) ? 1 : 0;
        if (ch.full) {
            P.tw = ch.full->tw;
            P.itw = ch.full->itw;
        }
        pm::BatchArgs ba{};
        std::vector<int> dec_key(n), crt_key(n);
        for (int k = 0; k < n; ++k) {
            const int i = order[i0 + k];
            const pm_plan* p = plans[i];
            ba.a[k] = a[i];
            ba.b[k] = b[i];
            ba.c[k] = c[i];
            ba.A[k] = p->A;
            ba.B[k] = p->B;
            ba.d[k] = p->desc;
            ba.K[k] = p->dev.K;
            dec_key[k] = p->dev.dec_ksteps;
            crt_key[k] = p->dev.crt_ksteps;
        }
        by_key(ba, n, dec_key, [&](const pm::BatchArgs& sub, int m, int ks) {
            pm::DevPlan Q = P;
            Q.dec_ksteps = ks;
            if (P.fused_cols)
                pm::launch_decompose_cols_batch(Q, sub, m, st);
            else
                pm::launch_decompose_mfma_batch(Q, sub, m, st);
        });
        if (P.fused_cols) {
            pm::launch_ntt_rows_fused_mid_batch(P, ba, n, st);
        } else {
            pm::launch_ntt_cols_batch(P, ba, n, true, st);
            pm::launch_ntt_rows_fused_batch(P, ba, n, st);
        }
        pm::launch_ntt_cols_batch(P, ba, n, false, st);
        by_key(ba, n, crt_key, [&](const pm::BatchArgs& sub, int m, int ks) {
            pm::DevPlan Q = P;
            Q.crt_ksteps = ks;
            pm::launch_crt_mfma_batch(Q, sub, m, st);
        });
    }
    check(hipGetLastError(), "kernel launch");
}

// The products split into sets whose members can share batched launches
// (first-fit over the products in order); a product that fits none (plain
// form, or a geometry without register kernels) is a set of its own, marked
// by `alone`.
struct ProductSets {
    std::vector<std::vector<int>> sets;
    std::vector<char> alone;
};

ProductSets product_sets(pm_plan* const* plans, int count) {
    ProductSets ps;
    for (int i = 0; i < count; ++i) {
        const bool ok = plans[i]->form == PM_FORM_MFMA && pm::batch_kernels_ok(plans[i]->dev);
        bool placed = false;
        for (size_t s = 0; ok && s < ps.sets.size() && !placed; ++s) {
            if (ps.alone[s]) continue;
            bool fits = true;
            for (int k : ps.sets[s]) fits = fits && plans[k] != plans[i] && groupable(plans[k], plans[i]);
            if (fits) {
                ps.sets[s].push_back(i);
                placed = true;
            }
        }
        if (!placed) {
            ps.sets.push_back({i});
            ps.alone.push_back(ok ? 0 : 1);
        }
    }
    return ps;
}

// Work of one product, for balancing runs and branches: K N (log N + 2) for
// the NTTs and pointwise steps, plus N L K / 8 for the contractions.
double product_work(const pm_plan* p) {
    const pm::HostPlan& h = p->host;
    return static_cast<double>(h.K) * h.N * (h.logN + 2) + static_cast<double>(h.N) * h.L * h.K / 8;
}

const char* const kPhaseNames[PM_NPHASES] = {"decompose_a", "decompose_b", "ntt_cols_fwd_a", "ntt_cols_fwd_b",
                                             "ntt_rows_fused", "ntt_cols_inv", "crt"};

}  // namespace

namespace pm {

// Phase i (0 .. PM_NPHASES - 1) of one product on device pointers, enqueued on
// st, with the residues of a and b in A and B (K x N words each).
// The product as one batched launch of its own (the kernels' per-product
// arguments for a batch of one).
BatchArgs one_product(pm_plan* pl, uint32_t* A, uint32_t* B, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    BatchArgs ba{};
    ba.a[0] = a;
    ba.b[0] = b;
    ba.c[0] = c;
    ba.A[0] = A;
    ba.B[0] = B;
    ba.d[0] = pl->desc;
    ba.K[0] = pl->dev.K;
    return ba;
}

// A plan whose product runs the fused column stages (matrix-core form only).
bool runs_fused_cols(const pm_plan* pl) { return pl->form == PM_FORM_MFMA && pl->dev.fused_cols; }

bool phase_empty(const pm_plan* pl, int i) { return runs_fused_cols(pl) && (i == 2 || i == 3); }

void run_phase_in(pm_plan* pl, uint32_t* A, uint32_t* B, const uint32_t* a, const uint32_t* b, uint32_t* c,
                  hipStream_t st, int i) {
    const pm::DevPlan& P = pl->dev;
    const bool mfma = pl->form == PM_FORM_MFMA;
    if (runs_fused_cols(pl) && i <= 4) {
        // decompose_a / _b include the first four column stages, the row pass
        // the next four; the forward column phases are empty
        const BatchArgs ba = one_product(pl, A, B, a, b, c);
        if (i <= 1) launch_decompose_cols_batch(P, ba, 1, st, i);
        if (i == 4) launch_ntt_rows_fused_mid_batch(P, ba, 1, st);
        return;
    }
    switch (i) {
        case 0: if (mfma) launch_decompose_mfma(P, a, A, st, true); else launch_decompose_valu(P, a, A, st); break;
        case 1:   // b scaled by the pointwise constant (b_side)
            if (mfma) launch_decompose_mfma(b_side(P), b, B, st, true); else launch_decompose_valu(b_side(P), b, B, st);
            break;
        case 2: launch_ntt_cols(P, A, true, st); break;
        case 3: launch_ntt_cols(P, B, true, st); break;
        case 4: launch_ntt_rows_fused(P, A, B, st); break;
        case 5: launch_ntt_cols(P, A, false, st); break;
        case 6: if (mfma) launch_crt_mfma(P, A, c, st); else launch_crt_valu(P, A, c, st); break;
        default: throw std::logic_error("no such phase");
    }
}

void run_phase(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st, int i) {
    run_phase_in(pl, pl->A, pl->B, a, b, c, st, i);
}

// The phases of one product on device pointers, enqueued on st, working in A
// and B. Where the register kernels apply, the decompositions of a and b
// (matrix-core form) and the forward column passes of A and B go out as one
// launch each: one pass alone can leave a partly filled last round of
// workgroups (8 K column blocks at N = 65536 against 1024 resident
// workgroups), which two passes in one launch share.
void run_device_in(pm_plan* pl, uint32_t* A, uint32_t* B, const uint32_t* a, const uint32_t* b, uint32_t* c,
                   hipStream_t st) {
    if (!batch_kernels_ok(pl->dev)) {
        for (int i = 0; i < PM_NPHASES; ++i) run_phase_in(pl, A, B, a, b, c, st, i);
    } else if (runs_fused_cols(pl)) {
        const BatchArgs ba = one_product(pl, A, B, a, b, c);
        launch_decompose_cols_batch(pl->dev, ba, 1, st);
        launch_ntt_rows_fused_mid_batch(pl->dev, ba, 1, st);
        for (int i = 5; i < PM_NPHASES; ++i) run_phase_in(pl, A, B, a, b, c, st, i);
    } else {
        const BatchArgs ba = one_product(pl, A, B, a, b, c);
        if (pl->form == PM_FORM_MFMA) {
            launch_decompose_mfma_batch(pl->dev, ba, 1, st);   // b with the b-side tables
        } else {
            run_phase_in(pl, A, B, a, b, c, st, 0);
            run_phase_in(pl, A, B, a, b, c, st, 1);
        }
        launch_ntt_cols_batch(pl->dev, ba, 1, true, st);
        for (int i = 4; i < PM_NPHASES; ++i) run_phase_in(pl, A, B, a, b, c, st, i);
    }
    check(hipGetLastError(), "kernel launch");
}

void run_device(pm_plan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c, hipStream_t st) {
    run_device_in(pl, pl->A, pl->B, a, b, c, st);
}

}  // namespace pm

using pm::run_device;

struct pm_batch {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    pm_batch_shape shape{};
    bool phased = false;
    double ms = 0;                                        // median replay of the timed choice, or 0
    std::vector<std::shared_ptr<pm::FullTables>> keep;    // full tables its table-twiddle runs read
    ~pm_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

// One chain of a captured batch: a batched run of products, or one product's
// own kernel chain.
struct Unit {
    std::vector<int> idx;
    bool batched;
    double work;
};

std::vector<Unit> batch_units(int count, pm_plan* const* plans, int runs) {
    const ProductSets ps = product_sets(plans, count);
    double total = 0;
    for (int i = 0; i < count; ++i) total += product_work(plans[i]);
    std::vector<Unit> units;
    for (size_t s = 0; s < ps.sets.size(); ++s) {
        std::vector<int> set = ps.sets[s];
        if (ps.alone[s] || runs <= 0) {
            for (int i : set) units.push_back({{i}, false, product_work(plans[i])});
            continue;
        }
        double sw = 0;
        for (int i : set) sw += product_work(plans[i]);
        const int r = std::max(1, std::min(static_cast<int>(set.size()), static_cast<int>(std::lround(runs * sw / total))));
        std::stable_sort(set.begin(), set.end(), [&](int x, int y) { return plans[x]->host.K > plans[y]->host.K; });
        // contiguous slices of about equal work
        Unit u{{}, true, 0.0};
        int cut = 0;
        double acc = 0;
        for (size_t k = 0; k < set.size(); ++k) {
            const double w = product_work(plans[set[k]]);
            u.idx.push_back(set[k]);
            u.work += w;
            acc += w;
            const size_t left = set.size() - k - 1;
            if (cut < r - 1 && left > 0 && (acc >= sw * (cut + 1) / r || left < static_cast<size_t>(r - 1 - cut))) {
                units.push_back(u);
                u = Unit{{}, true, 0.0};
                ++cut;
            }
        }
        if (!u.idx.empty()) units.push_back(u);
    }
    return units;
}

// Captures the batch under `shape` into bt's graph; ntt_form and twiddles of -1
// keep the plans' own.
void capture_batch(pm_batch& bt, int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                   uint32_t* const* c, const pm_batch_shape& shape) {
    need(shape.runs >= 0 && shape.streams >= 1, "shape: runs must be >= 0 and streams >= 1");
    need(shape.ntt_form == -1 || shape.ntt_form == PM_NTT_VALU || shape.ntt_form == PM_NTT_MFMA,
         "shape: ntt_form must be -1, PM_NTT_VALU or PM_NTT_MFMA");
    need(shape.twiddles == -1 || shape.twiddles == PM_TWIDDLES_TABLE || shape.twiddles == PM_TWIDDLES_GEN,
         "shape: twiddles must be -1, PM_TWIDDLES_TABLE or PM_TWIDDLES_GEN");
    std::vector<Unit> units = batch_units(count, plans, shape.runs);
    std::stable_sort(units.begin(), units.end(), [](const Unit& x, const Unit& y) { return x.work > y.work; });
    // per batched run: its NTT choice (the matrix-core form where the run's geometry has it)
    std::vector<NttChoice> choice(units.size());
    std::vector<std::shared_ptr<pm::FullTables>> keep;
    bool phased = false;
    for (size_t u = 0; u < units.size(); ++u) {
        if (!units[u].batched) continue;
        phased = true;
        pm_plan* top = plans[units[u].idx[0]];   // most primes: the run's sort order
        for (int i : units[u].idx)
            if (plans[i]->host.K > top->host.K) top = plans[i];
        NttChoice& ch = choice[u];
        if (shape.ntt_form >= 0) ch.col_mfma = shape.ntt_form == PM_NTT_MFMA && top->dev.colm_fwd && top->dev.colm_inv;
        if (shape.twiddles == PM_TWIDDLES_GEN) ch.row_gen = 1;
        if (shape.twiddles == PM_TWIDDLES_TABLE) {
            ch.row_gen = 0;
            auto f = top->full ? top->full : pm::full_tables(*top->tw, top->host);
            ch.full = f.get();
            keep.push_back(f);
        }
    }
    const size_t nb = std::max<size_t>(1, std::min(units.size(), static_cast<size_t>(shape.streams)));
    hipStream_t origin = nullptr;
    check(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking), "hipStreamCreate");
    std::vector<hipStream_t> branch(nb, nullptr);
    std::vector<hipEvent_t> done(nb, nullptr);
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&] {
        for (auto e : done)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto st : branch)
            if (st) (void)hipStreamDestroy(st);
        (void)hipStreamDestroy(origin);
    };
    try {
        for (auto& st : branch) check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : done) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreate");
        check(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            check(hipEventRecord(fork, origin), "hipEventRecord");
            for (size_t k = 0; k < nb; ++k) check(hipStreamWaitEvent(branch[k], fork, 0), "wait");
            // largest work first, each onto the branch with the least work so far
            std::vector<double> load(nb, 0.0);
            for (size_t u = 0; u < units.size(); ++u) {
                const size_t k = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
                if (units[u].batched) {
                    run_batched(units[u].idx, plans, a, b, c, branch[k], choice[u]);
                } else {
                    const int i = units[u].idx[0];
                    run_device(plans[i], a[i], b[i], c[i], branch[k]);
                }
                load[k] += units[u].work;
            }
            for (size_t k = 0; k < nb; ++k) {
                check(hipEventRecord(done[k], branch[k]), "hipEventRecord");
                check(hipStreamWaitEvent(origin, done[k], 0), "wait");
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(origin, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(origin, &graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    } catch (...) {
        if (graph) (void)hipGraphDestroy(graph);
        cleanup();
        throw;
    }
    cleanup();
    if (bt.exec) (void)hipGraphExecDestroy(bt.exec);
    if (bt.graph) (void)hipGraphDestroy(bt.graph);
    bt.graph = graph;
    bt.exec = exec;
    bt.shape = shape;
    bt.shape.streams = static_cast<int>(nb);
    bt.phased = phased;
    bt.keep = std::move(keep);
}

// Median milliseconds of one replay of bt's graph: after one warm-up,
// `samples` samples of `reps` back-to-back replays each.
double time_batch(const pm_batch& bt, hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps = 2, int samples = 5) {
    check(hipGraphLaunch(bt.exec, st), "hipGraphLaunch");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::vector<double> t;
    for (int r = 0; r < samples; ++r) {
        check(hipEventRecord(e0, st), "hipEventRecord");
        for (int k = 0; k < reps; ++k) check(hipGraphLaunch(bt.exec, st), "hipGraphLaunch");
        check(hipEventRecord(e1, st), "hipEventRecord");
        check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        t.push_back(ms / reps);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

int env_int(const char* name, int fallback) {
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    char* end = nullptr;
    const long x = std::strtol(v, &end, 10);
    need(end && *end == 0 && x >= 0 && x <= 4096, (std::string(name) + " must be a non-negative integer").c_str());
    return static_cast<int>(x);
}

// pm_batch_create's timed choice, dimension by dimension: (runs, streams) with
// the plans' own forms, then the column-pass form, then the twiddle source;
// a later candidate replaces the best so far only if at least 1 % faster.
void tune_batch(pm_batch& bt, int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                uint32_t* const* c) {
    const int runs_env = std::getenv("PM_BATCH_RUNS") ? env_int("PM_BATCH_RUNS", 0) : env_int("PM_BATCH_GROUPS", -1);
    const int streams_env =
        std::getenv("PM_BATCH_STREAMS") ? env_int("PM_BATCH_STREAMS", 1) : env_int("PM_BATCH_BRANCHES", -1);
    const bool tune = env_int("PM_BATCH_TUNE", 1) != 0;
    if (!tune) {
        pm_batch_shape sh{runs_env >= 0 ? runs_env : 16, streams_env >= 1 ? streams_env : 16, -1, -1};
        capture_batch(bt, count, plans, a, b, c, sh);
        return;
    }
    const ProductSets ps = product_sets(plans, count);
    int groups = 0, grouped = 0;
    for (size_t s = 0; s < ps.sets.size(); ++s)
        if (!ps.alone[s]) {
            ++groups;
            grouped += static_cast<int>(ps.sets[s].size());
        }
    std::vector<pm_batch_shape> cands;
    auto add = [&](int r, int st) {
        if (runs_env >= 0) r = runs_env;
        if (streams_env >= 1) st = streams_env;
        for (const auto& x : cands)
            if (x.runs == r && x.streams == st) return;
        cands.push_back(pm_batch_shape{r, st, -1, -1});
    };
    if (groups > 0) {
        for (int m : {16, 8, 4, 2, 1}) {
            const int r = std::min(grouped, m * groups);
            add(r, std::min(r + count - grouped, 16));
            add(r, std::min(r + count - grouped, 4));
        }
    }
    add(0, 4);
    add(0, 16);
    // Timed on the null stream, which lives as long as the process: a graph's
    // executable keeps state about the stream of its last launch, so the
    // stream that launched the chosen graph must outlive it.
    const hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        pm_batch best;
        double best_ms = 0;
        std::vector<std::pair<double, pm_batch_shape>> timed;   // every (runs, streams) candidate
        auto take = [&](pm_batch& trial, double ms) {
            std::swap(best.graph, trial.graph);
            std::swap(best.exec, trial.exec);
            best.shape = trial.shape;
            best.phased = trial.phased;
            best.keep.swap(trial.keep);
            best_ms = ms;
        };
        auto consider = [&](const pm_batch_shape& sh) {
            pm_batch trial;
            capture_batch(trial, count, plans, a, b, c, sh);
            const double ms = time_batch(trial, st, e0, e1);
            timed.emplace_back(ms, trial.shape);
            if (!best.exec || ms < 0.99 * best_ms) take(trial, ms);
        };
        for (const auto& sh : cands) consider(sh);
        // Short samples on a GPU that idles between candidates run at a higher
        // clock than a sustained run of replays, and not by the same amount for
        // every shape: the three fastest meet again in alternation, three
        // rounds of ten back-to-back replays each, and the mean decides.
        std::sort(timed.begin(), timed.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        if (timed.size() > 1) {
            const size_t n = std::min<size_t>(3, timed.size());
            std::vector<std::unique_ptr<pm_batch>> fin(n);
            std::vector<double> tot(n, 0.0);
            for (size_t i = 0; i < n; ++i) {
                fin[i].reset(new pm_batch);
                capture_batch(*fin[i], count, plans, a, b, c, timed[i].second);
            }
            for (int r = 0; r < 3; ++r)
                for (size_t i = 0; i < n; ++i) tot[i] += time_batch(*fin[i], st, e0, e1, 10, 1);
            size_t w = 0;
            for (size_t i = 1; i < n; ++i)
                if (tot[i] < tot[w]) w = i;
            take(*fin[w], 0.0);
            // the forms below are timed as the shapes were, against the same kind of sample
            best_ms = time_batch(best, st, e0, e1);
        }
        if (best.phased && !std::getenv("PM_NTT_FORM")) {
            bool mfma_tables = false;
            for (int i = 0; i < count; ++i) mfma_tables = mfma_tables || (plans[i]->dev.colm_fwd && plans[i]->dev.colm_inv);
            pm_batch_shape sh = best.shape;
            sh.ntt_form = PM_NTT_VALU;
            consider(sh);
            if (mfma_tables) {
                sh.ntt_form = PM_NTT_MFMA;
                consider(sh);
            }
        }
        if (best.phased && !std::getenv("FHERMA_TWIDDLES")) {
            pm_batch_shape sh = best.shape;
            sh.twiddles = PM_TWIDDLES_GEN;
            consider(sh);
            sh.twiddles = PM_TWIDDLES_TABLE;
            consider(sh);
        }
        std::swap(bt.graph, best.graph);
        std::swap(bt.exec, best.exec);
        bt.shape = best.shape;
        bt.phased = best.phased;
        bt.keep.swap(best.keep);
        bt.ms = best_ms;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// The timed choice of a plan's row-pass twiddle source and column-pass form:
// device-resident products on dummy operands (zeros), two warm-up runs and the
// median of five per candidate; the plan's own setting (the reference's size
// rule for the twiddles, the butterflies for the columns) is kept unless
// another candidate is at least 5 % faster, the reference's rule for its host
// modes (solve.cu:817-862). FHERMA_TWIDDLES / PM_NTT_FORM in the environment
// pin their dimension. Returns the chosen setting's median ms per product.
double tune_kernels(pm_plan* pl) {
    const size_t words = pl->words();
    uint32_t* buf = static_cast<uint32_t*>(dev_alloc(3 * words * sizeof(uint32_t)));
    double best = 0;
    try {
        check(hipMemset(buf, 0, 3 * words * sizeof(uint32_t)), "hipMemset");
        const uint32_t *a = buf, *b = buf + words;
        uint32_t* c = buf + 2 * words;
        const bool log = std::getenv("FHERMA_AUTOTUNE_LOG") || std::getenv("PM_HOST_TUNE_LOG");
        auto time_it = [&] {
            std::vector<double> t;
            for (int r = 0; r < 7; ++r) {
                check(hipEventRecord(pl->ev[0], pl->stream), "hipEventRecord");
                run_device(pl, a, b, c, pl->stream);
                check(hipEventRecord(pl->ev[1], pl->stream), "hipEventRecord");
                check(hipEventSynchronize(pl->ev[1]), "hipEventSynchronize");
                float ms = 0;
                check(hipEventElapsedTime(&ms, pl->ev[0], pl->ev[1]), "hipEventElapsedTime");
                if (r >= 2) t.push_back(ms);
            }
            std::sort(t.begin(), t.end());
            if (log)
                std::fprintf(stderr, "device twiddles %s, column pass %s: %.3f ms (median of 5)\n",
                             pl->dev.row_gen ? "gen" : "table", pl->dev.col_mfma ? "mfma" : "valu", t[2]);
            return t[2];
        };
        best = time_it();
        if (!std::getenv("FHERMA_TWIDDLES") && pm::batch_kernels_ok(pl->dev)) {
            const int own = pl->dev.row_gen ? PM_TWIDDLES_GEN : PM_TWIDDLES_TABLE;
            const int other = own == PM_TWIDDLES_GEN ? PM_TWIDDLES_TABLE : PM_TWIDDLES_GEN;
            if (pm_plan_set_twiddles(pl, other) != PM_OK) throw std::runtime_error(g_err);
            const double ms = time_it();
            if (ms < 0.95 * best)
                best = ms;
            else if (pm_plan_set_twiddles(pl, own) != PM_OK)
                throw std::runtime_error(g_err);
        }
        if (!std::getenv("PM_NTT_FORM") && pl->dev.colm_fwd && pl->dev.colm_inv) {
            const int own = pl->dev.col_mfma;
            pl->dev.col_mfma = own ? 0 : 1;
            const double ms = time_it();
            if (ms < 0.95 * best)
                best = ms;
            else
                pl->dev.col_mfma = own;
        }
    } catch (...) {
        (void)hipFree(buf);
        throw;
    }
    check(hipFree(buf), "hipFree");
    return best;
}

void check_batch_args(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                      uint32_t* const* c, pm_batch** out) {
    need(out != nullptr && count > 0 && plans && a && b && c, "bad argument");
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
        need(plans[i] && a[i] && b[i] && c[i], "null entry in batch");
        for (int k = 0; k < i; ++k) need(plans[k] != plans[i], "each product of a batch needs its own plan");
    }
}

}  // namespace

extern "C" {

const char* pm_last_error(void) { return g_err.c_str(); }

int pm_plan_create(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, pm_plan** out) {
    return guarded([&] {
        need(out != nullptr, "out is null");
        *out = nullptr;
        *out = make_plan_object(N, W, L, q);
    });
}

void pm_plan_destroy(pm_plan* plan) { delete plan; }

int pm_plan_query(const pm_plan* plan, uint32_t* N, uint32_t* L, uint32_t* K) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        if (N) *N = plan->host.N;
        if (L) *L = static_cast<uint32_t>(plan->host.L);
        if (K) *K = static_cast<uint32_t>(plan->host.K);
    });
}

int pm_plan_primes(const pm_plan* plan, uint32_t* out, uint32_t cap) {
    return guarded([&] {
        need(plan != nullptr && out != nullptr, "null argument");
        need(cap >= static_cast<uint32_t>(plan->host.K), "cap is smaller than K");
        std::memcpy(out, plan->host.p.data(), plan->host.p.size() * sizeof(uint32_t));
    });
}

int pm_choose_primes(uint32_t N, uint32_t qbits, uint32_t* out, uint32_t cap, uint32_t* K) {
    return guarded([&] {
        need(out != nullptr && K != nullptr, "null argument");
        need(N >= 2 && (N & (N - 1)) == 0 && N <= (1u << pm::kMaxLogN), "N must be a power of two in [2, 2^22]");
        const std::vector<uint32_t> p = pm::choose_primes(N, static_cast<int>(qbits));
        need(p.size() <= cap, "cap is smaller than K");
        std::memcpy(out, p.data(), p.size() * sizeof(uint32_t));
        *K = static_cast<uint32_t>(p.size());
    });
}

int pm_plan_set_twiddles(pm_plan* plan, int source) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        need(source == PM_TWIDDLES_TABLE || source == PM_TWIDDLES_GEN,
             "source must be PM_TWIDDLES_TABLE or PM_TWIDDLES_GEN");
        if (source == PM_TWIDDLES_TABLE && !plan->full) plan->full = pm::full_tables(*plan->tw, plan->host);
        plan->dev.row_gen = source == PM_TWIDDLES_GEN ? 1 : 0;
        if (source == PM_TWIDDLES_GEN && pm::batch_kernels_ok(plan->dev)) {
            // nothing of this plan reads the full tables now (hipFree waits for
            // work in flight, should this release the last copy)
            plan->full.reset();
        }
        plan->dev.tw = plan->full ? plan->full->tw : nullptr;
        plan->dev.itw = plan->full ? plan->full->itw : nullptr;
    });
}

int pm_plan_set_ntt_form(pm_plan* plan, int form) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        need(form == PM_NTT_VALU || form == PM_NTT_MFMA, "form must be PM_NTT_VALU or PM_NTT_MFMA");
        need(form == PM_NTT_VALU || (plan->dev.colm_fwd && plan->dev.colm_inv),
             "the matrix-core column pass needs R = 256 rows (N = 65536)");
        plan->dev.col_mfma = form == PM_NTT_MFMA ? 1 : 0;
    });
}

int pm_plan_set_fused_cols(pm_plan* plan, int on) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        need(on == 0 || on == 1, "fused_cols must be 0 or 1");
        need(on == 0 || pm::fused_cols_ok(plan->dev), "fused column stages need N = 32768 or 65536 and the matrix-core tables");
        plan->dev.fused_cols = on;
    });
}

int pm_plan_get_fused_cols(const pm_plan* plan) { return plan ? plan->dev.fused_cols : -1; }

int pm_plan_get_ntt_form(const pm_plan* plan) { return plan ? (plan->dev.col_mfma ? PM_NTT_MFMA : PM_NTT_VALU) : -1; }

int pm_plan_get_twiddles(const pm_plan* plan) {
    return plan ? (plan->dev.row_gen ? PM_TWIDDLES_GEN : PM_TWIDDLES_TABLE) : -1;
}

int pm_plan_set_form(pm_plan* plan, int form) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        need(form == PM_FORM_PLAIN || form == PM_FORM_MFMA, "form must be PM_FORM_PLAIN or PM_FORM_MFMA");
        if (form == PM_FORM_MFMA) plan->ensure_mfma();
        plan->form = form;
    });
}

int pm_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream) {
    return guarded([&] {
        need(plan && a && b && c, "null argument");
        run_device(plan, a, b, c, stream_of(plan, stream));
    });
}

int pm_multiply_device_noalloc(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream) {
    return guarded([&] {
        need(plan && a && b && c, "null argument");
        // A, then B at the next 256-byte boundary (as separate allocations would be)
        const size_t kn = plan->residues(), b_at = (kn + 63) / 64 * 64;
        if (b_at + kn > pm::kScratchWords) {
            throw std::invalid_argument("pm_multiply_device_noalloc: the product needs " + std::to_string(b_at + kn) +
                                        " words of residue workspace (2 K N); the static workspace holds " +
                                        std::to_string(pm::kScratchWords));
        }
        uint32_t* A = pm::scratch_on_current_device();
        pm::run_device_in(plan, A, A + b_at, a, b, c, stream_of(plan, stream));
    });
}

int pm_multiply(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    return guarded([&] {
        need(plan && a && b && c, "null argument");
        pm::host_multiply(plan, a, b, c);
    });
}

int pm_batch_multiply_host(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                           uint32_t* const* c, int depth) {
    return guarded([&] {
        need(count >= 0 && (count == 0 || (plans && a && b && c)), "null argument");
        for (int i = 0; i < count; ++i) need(plans[i] && a[i] && b[i] && c[i], "null entry in batch");
        if (depth <= 0) depth = env_int("PM_HOST_DEPTH", 3);
        need(depth >= 1 && depth <= 16, "depth must be 1 .. 16");
        pm::host_multiply_batch(count, plans, a, b, c, depth);
    });
}

int pm_plan_set_host_mode(pm_plan* plan, int mode) { return pm_plan_set_host_modes(plan, mode, mode); }

int pm_plan_set_host_modes(pm_plan* plan, int in_mode, int out_mode) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        need(in_mode >= PM_HOST_STAGED && in_mode <= PM_HOST_DIRECT && out_mode >= PM_HOST_STAGED &&
                 out_mode <= PM_HOST_DIRECT,
             "unknown host mode");
        pm::host_set_modes(plan, in_mode, out_mode);
    });
}

int pm_plan_get_host_modes(const pm_plan* plan, int* in_mode, int* out_mode) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        if (in_mode) *in_mode = pm::host_get_mode(plan);
        if (out_mode) *out_mode = pm::host_get_out_mode(plan);
    });
}

int pm_plan_tune_host(pm_plan* plan, int* mode, double* ms_per_product) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        const double ms = pm::host_autotune(plan);
        if (mode) *mode = pm::host_get_mode(plan);
        if (ms_per_product) *ms_per_product = ms;
    });
}

int pm_batch_multiply_device(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                             uint32_t* const* c, void* stream) {
    return guarded([&] {
        need(count >= 0 && (count == 0 || (plans && a && b && c)), "null argument");
        for (int i = 0; i < count; ++i) need(plans[i] && a[i] && b[i] && c[i], "null entry in batch");
        if (!stream) {
            for (int i = 0; i < count; ++i) run_device(plans[i], a[i], b[i], c[i], stream_of(plans[i], stream));
            return;
        }
        const hipStream_t st = static_cast<hipStream_t>(stream);
        const ProductSets ps = product_sets(plans, count);
        for (size_t s = 0; s < ps.sets.size(); ++s) {
            if (ps.alone[s]) {
                for (int i : ps.sets[s]) run_device(plans[i], a[i], b[i], c[i], st);
            } else {
                run_batched(ps.sets[s], plans, a, b, c, st);
            }
        }
    });
}

int pm_batch_create(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                    uint32_t* const* c, pm_batch** out) {
    return guarded([&] {
        check_batch_args(count, plans, a, b, c, out);
        std::unique_ptr<pm_batch> bt(new pm_batch());
        tune_batch(*bt, count, plans, a, b, c);
        *out = bt.release();
    });
}

int pm_batch_create_shaped(int count, pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                           uint32_t* const* c, const pm_batch_shape* shape, pm_batch** out) {
    return guarded([&] {
        check_batch_args(count, plans, a, b, c, out);
        need(shape != nullptr, "shape is null");
        std::unique_ptr<pm_batch> bt(new pm_batch());
        capture_batch(*bt, count, plans, a, b, c, *shape);
        *out = bt.release();
    });
}

int pm_batch_launch(pm_batch* batch, void* stream) {
    return guarded([&] {
        need(batch != nullptr, "batch is null");
        check(hipGraphLaunch(batch->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
    });
}

int pm_batch_get_shape(const pm_batch* batch, pm_batch_shape* shape, double* ms) {
    return guarded([&] {
        need(batch != nullptr, "batch is null");
        if (shape) *shape = batch->shape;
        if (ms) *ms = batch->ms;
    });
}

int pm_batch_phased(const pm_batch* batch) { return batch && batch->phased ? 1 : 0; }

void pm_batch_destroy(pm_batch* batch) { delete batch; }

int pm_stage_decompose(pm_plan* plan, const uint32_t* limbs, uint32_t* res, void* stream) {
    return guarded([&] {
        need(plan && limbs && res, "null argument");
        if (plan->form == PM_FORM_MFMA)
            pm::launch_decompose_mfma(plan->dev, limbs, res, stream_of(plan, stream));
        else
            pm::launch_decompose_valu(plan->dev, limbs, res, stream_of(plan, stream));
        check(hipGetLastError(), "kernel launch");
    });
}

int pm_stage_ntt(pm_plan* plan, uint32_t* res, int forward, void* stream) {
    return guarded([&] {
        need(plan && res, "null argument");
        hipStream_t st = stream_of(plan, stream);
        if (forward) {
            pm::launch_ntt_cols(plan->dev, res, true, st);
            pm::launch_ntt_rows(plan->dev, res, true, st);
        } else {
            pm::launch_ntt_rows(plan->dev, res, false, st);
            pm::launch_ntt_cols(plan->dev, res, false, st);
        }
        check(hipGetLastError(), "kernel launch");
    });
}

int pm_stage_pointwise(pm_plan* plan, uint32_t* A, const uint32_t* B, void* stream) {
    return guarded([&] {
        need(plan && A && B, "null argument");
        pm::launch_pointwise(plan->dev, A, B, stream_of(plan, stream));
        check(hipGetLastError(), "kernel launch");
    });
}

int pm_stage_crt(pm_plan* plan, const uint32_t* res, uint32_t* c, void* stream) {
    return guarded([&] {
        need(plan && res && c, "null argument");
        if (plan->form == PM_FORM_MFMA)
            pm::launch_crt_mfma(plan->dev, res, c, stream_of(plan, stream));
        else
            pm::launch_crt_valu(plan->dev, res, c, stream_of(plan, stream));
        check(hipGetLastError(), "kernel launch");
    });
}

int pm_time_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, int warmup,
                            int iters, double* ms_per_product, void* stream) {
    return guarded([&] {
        need(plan && a && b && c && ms_per_product && iters > 0 && warmup >= 0, "bad argument");
        hipStream_t st = stream_of(plan, stream);
        for (int i = 0; i < warmup; ++i) run_device(plan, a, b, c, st);
        check(hipEventRecord(plan->ev[0], st), "hipEventRecord");
        for (int i = 0; i < iters; ++i) run_device(plan, a, b, c, st);
        check(hipEventRecord(plan->ev[1], st), "hipEventRecord");
        check(hipEventSynchronize(plan->ev[1]), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, plan->ev[0], plan->ev[1]), "hipEventElapsedTime");
        *ms_per_product = static_cast<double>(ms) / iters;
    });
}

int pm_profile_phases(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, int iters,
                      double* ms_per_phase, void* stream) {
    return guarded([&] {
        need(plan && a && b && c && ms_per_phase && iters > 0, "bad argument");
        hipStream_t st = stream_of(plan, stream);
        // Each phase's kernel runs `iters` times back to back between two events
        // (re-running a phase on its own output is well defined: every kernel
        // maps residues to residues), so the mean includes one dispatch gap and
        // no host latency. A last full product leaves the true result in c.
        for (int i = 0; i < PM_NPHASES; ++i) {
            check(hipEventRecord(plan->ev[i], st), "hipEventRecord");
            for (int it = 0; it < iters; ++it) pm::run_phase(plan, a, b, c, st, i);
            check(hipEventRecord(plan->ev_stop[i], st), "hipEventRecord");
        }
        check(hipGetLastError(), "kernel launch");
        run_device(plan, a, b, c, st);
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        for (int i = 0; i < PM_NPHASES; ++i) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, plan->ev[i], plan->ev_stop[i]), "hipEventElapsedTime");
            // the forward column phases of a plan with fused column stages launch nothing
            ms_per_phase[i] = pm::phase_empty(plan, i) ? 0.0 : ms / iters;
        }
    });
}

const char* pm_phase_name(int phase) { return phase >= 0 && phase < PM_NPHASES ? kPhaseNames[phase] : ""; }

int pm_device_alloc(size_t bytes, void** out) {
    return guarded([&] {
        need(out != nullptr, "out is null");
        *out = dev_alloc(bytes);
    });
}

int pm_device_free(void* ptr) {
    return guarded([&] {
        if (ptr) check(hipFree(ptr), "hipFree");
    });
}

int pm_copy_to_device(void* dst, const void* src, size_t bytes) {
    return guarded([&] { check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D"); });
}

int pm_copy_to_host(void* dst, const void* src, size_t bytes) {
    return guarded([&] { check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H"); });
}

int pm_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        need(dst && src, "null pointer");
        need(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(src) % 16 == 0,
             "pm_device_copy needs 16-byte aligned pointers and sizes");
        pm::launch_copy16(dst, src, bytes, static_cast<hipStream_t>(stream));
        check(hipGetLastError(), "kernel launch");
    });
}

int pm_device_sum(const void* src, size_t bytes, uint64_t* sum) {
    return guarded([&] {
        need(src && sum, "null pointer");
        need(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0,
             "pm_device_sum needs a 16-byte aligned pointer and size");
        uint64_t* d = static_cast<uint64_t*>(dev_alloc((pm::kSumBlocks + 1) * sizeof(uint64_t)));
        try {
            pm::launch_sum16(d, d + pm::kSumBlocks, src, bytes, nullptr);
            check(hipGetLastError(), "kernel launch");
            check(hipMemcpy(sum, d + pm::kSumBlocks, sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        check(hipFree(d), "hipFree");
    });
}

int pm_stream_create(void** out) {
    return guarded([&] {
        need(out != nullptr, "out is null");
        hipStream_t s = nullptr;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *out = s;
    });
}

int pm_stream_destroy(void* stream) {
    return guarded([&] {
        if (stream) check(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy");
    });
}

int pm_stream_synchronize(void* stream) {
    return guarded([&] { check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize"); });
}

int pm_device_synchronize(void) {
    return guarded([&] { check(hipDeviceSynchronize(), "hipDeviceSynchronize"); });
}

int pm_probe_issue(int kind, int iters, double* ms, double* valu_per_simd, double* mfma_per_simd, double* clock_mhz) {
    return guarded([&] {
        need(ms && valu_per_simd && mfma_per_simd && clock_mhz && iters > 0, "bad argument");
        pm::probe_issue(kind, iters, ms, valu_per_simd, mfma_per_simd, clock_mhz);
    });
}

int pm_device_info(char* name, size_t len, int* compute_units) {
    return guarded([&] {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        if (name && len) {
            std::strncpy(name, prop.gcnArchName, len - 1);
            name[len - 1] = 0;
        }
        if (compute_units) *compute_units = prop.multiProcessorCount;
    });
}

}  // extern "C"

// --- FHERMA C++ boundary ------------------------------------------------------

int pm_plan_tune_kernels(pm_plan* plan, double* ms_per_product) {
    return guarded([&] {
        need(plan != nullptr, "plan is null");
        const double ms = tune_kernels(plan);
        if (ms_per_product) *ms_per_product = ms;
    });
}

void* fherma_init(const fherma::Point& p) {
    if (p.q.data.size() != p.L) throw std::runtime_error("q must have L limbs");
    pm_plan* plan = nullptr;
    if (pm_plan_create(p.N, p.W, p.L, p.q.data.data(), &plan) != PM_OK) throw std::runtime_error(g_err);
    // Like the reference's init (solve.cu:710-869): the host knobs of the
    // environment, then, unless they name a mode or FHERMA_AUTOTUNE=0, the timed
    // choice of the host modes on dummy operands (solve.cu:817-862).
    const int rc = guarded([&] {
        if (pm::host_configure_from_env(plan)) {
            pm::host_autotune(plan);
            (void)tune_kernels(plan);
        }
    });
    if (rc != PM_OK) {
        const std::string msg = g_err;
        pm_plan_destroy(plan);
        throw std::runtime_error(msg);
    }
    return plan;
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* plan = static_cast<pm_plan*>(state);
    if (!plan) throw std::runtime_error("fherma_run: null state");
    const size_t words = plan->words();
    if (in.a.data.size() != words || in.b.data.size() != words) {
        throw std::runtime_error("operand size does not match the point");
    }
    fherma::Outputs out;
    out.c.shape = {static_cast<int64_t>(plan->host.N), static_cast<int64_t>(plan->host.L)};
    // The answer is sized (and zero-filled, as std::vector does) on the host
    // pipeline's helper thread while the inputs stream (solve.cu:907-918).
    const int rc = guarded([&] {
        pm::host_multiply_sized(plan, in.a.data.data(), in.b.data.data(), [&out, words] { out.c.data.resize(words); },
                                [&out] { return out.c.data.data(); });
    });
    if (rc != PM_OK) throw std::runtime_error(g_err);
    return out;
}

void fherma_free(void* state) { pm_plan_destroy(static_cast<pm_plan*>(state)); }
