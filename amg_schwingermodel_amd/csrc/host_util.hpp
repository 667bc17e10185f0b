// Host-only helpers shared by the C ABI implementation.
#pragma once
#include <cstddef>
#include <cstdint>

namespace amgs {

// The reference's test-vector random stream (src/amg.cpp:8-35).
class TvStream {
public:
    explicit TvStream(uint32_t seed);
    ~TvStream();
    TvStream(const TvStream&) = delete;
    TvStream& operator=(const TvStream&) = delete;
    double uniform();
    // ncomplex interleaved complex entries, imaginary part drawn first
    void fill(double* out, size_t ncomplex);

private:
    struct Impl;
    Impl* p;
};

// true if no entry of p[0..n) is NaN or infinite (OpenMP over the entries)
bool all_finite(const double* p, size_t n);
// memcpy over the host's threads (1 MiB pieces)
void par_copy(void* dst, const void* src, size_t bytes);

}  // namespace amgs
