// The FHERMA polymul/negacyclic@1.0.0 plugin boundary: the wire types and the
// three entry points a platform harness binds by name.
//
// These declarations must stay layout-identical to the spec's generated header
// (reference fherma.h:13-48) so that an unmodified harness links against this
// library: the same namespace, struct members in the same order, std::vector
// storage, and C++ (not C) linkage for the three functions.
//
//   void* fherma_init(const fherma::Point&)            reference fherma.h:46, solve.cu:710
//   fherma::Outputs fherma_run(void*, const Inputs&)   reference fherma.h:47, solve.cu:871
//   void fherma_free(void*)                            reference fherma.h:48, solve.cu:1061
//
// Errors are thrown as std::runtime_error, as the reference does.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fherma {

template <class T>
struct Tensor {
    std::vector<int64_t> shape;
    std::vector<T> data;   // row-major

    int64_t count() const {
        int64_t total = 1;
        for (auto d : shape) total *= d;
        return total;
    }
};

struct Point {
    uint32_t N = 0;        // ring degree, a power of two in [2, 2^22]
    uint32_t W = 0;        // bit width of q
    uint32_t L = 0;        // u32 limbs per coefficient, ceil(W / 32), at most 64
    Tensor<uint32_t> q;    // L limbs, little-endian
};

struct Inputs {
    Tensor<uint32_t> a;    // N x L
    Tensor<uint32_t> b;    // N x L
};

struct Outputs {
    Tensor<uint32_t> c;    // N x L: a * b mod (X^N + 1, q), canonical in [0, q)
};

}  // namespace fherma

void* fherma_init(const fherma::Point& p);
fherma::Outputs fherma_run(void* state, const fherma::Inputs& in);
void fherma_free(void* state);
