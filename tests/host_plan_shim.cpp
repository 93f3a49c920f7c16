// Test shim: the library's host plan (fherma-polymult_amd/csrc/pm_plan.cpp),
// compiled with g++ alone, exported with C linkage so that tests/test_host_plan.py
// can compare its constants with the oracle's on the CPU.
#include "pm_plan.h"

#include <cstring>
#include <stdexcept>

extern "C" {

// Fills the plan's constants for (N, W, q[L]); returns K, or -1 with the
// exception's message in err.
int hp_plan(uint32_t N, uint32_t W, const uint32_t* q, int L, uint32_t* primes, uint32_t* pneg, uint32_t* psi,
            uint32_t* ipsi, uint32_t* scale, uint32_t* pow2, uint32_t* pow2b, uint32_t* M, uint32_t* negvP, double* mq,
            double* nq, char* err, int errlen) {
    try {
        const pm::HostPlan h = pm::make_plan(N, W, q, L);
        const int K = h.K;
        std::memcpy(primes, h.p.data(), 4 * K);
        std::memcpy(pneg, h.pneg.data(), 4 * K);
        std::memcpy(psi, h.psi.data(), 4 * K);
        std::memcpy(ipsi, h.ipsi.data(), 4 * K);
        std::memcpy(scale, h.scale.data(), 4 * K);
        std::memcpy(pow2, h.pow2.data(), 4 * K * L);
        std::memcpy(pow2b, h.pow2b.data(), 4 * K * L);
        std::memcpy(M, h.M.data(), 4 * K * L);
        std::memcpy(negvP, h.negvP.data(), 4 * (K + 1) * L);
        std::memcpy(mq, h.mq.data(), 8 * K);
        std::memcpy(nq, h.nq.data(), 8 * (K + 1));
        return K;
    } catch (const std::exception& e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return -1;
    }
}

}  // extern "C"
