// TEST INFRASTRUCTURE ONLY: a C-callable face for the reference's own CPU path.
//
// oracle/Makefile compiles this file together with the reference's
// solve_cpu.cpp (which brings rns.cuh, rns_kernels.cuh, rns_plan.h and fherma.h
// from the reference directory) into oracle/_ref/libref_solve_cpu.so. The
// reference sources are compiled where they lie and never copied. The three
// entry points it calls are the reference's plugin boundary
// (fherma.h:46-48; solve_cpu.cpp:204-249).
#include "fherma.h"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

namespace {
std::string g_err;
}

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Returns the reference's opaque state, or null with ref_last_error() set.
void* ref_init(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q) {
    try {
        fherma::Point p{};
        p.N = N;
        p.W = W;
        p.L = L;
        p.q.shape = {static_cast<int64_t>(L)};
        p.q.data.assign(q, q + L);
        return fherma_init(p);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// c = a * b mod (X^N + 1, q) through the reference's fherma_run; 0 on success.
int ref_run(void* state, uint32_t N, uint32_t L, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    try {
        const size_t words = static_cast<size_t>(N) * L;
        fherma::Inputs in{};
        in.a.shape = in.b.shape = {static_cast<int64_t>(N), static_cast<int64_t>(L)};
        in.a.data.assign(a, a + words);
        in.b.data.assign(b, b + words);
        fherma::Outputs out = fherma_run(state, in);
        if (out.c.data.size() != words) {
            g_err = "reference returned the wrong output size";
            return -1;
        }
        std::memcpy(c, out.c.data.data(), words * sizeof(uint32_t));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_free(void* state) { fherma_free(state); }

}  // extern "C"
