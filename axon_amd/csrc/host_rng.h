// Host-side random streams of the reference protocol.
//
// The reference draws its stimuli on the host with std::mt19937 (mm:793,
// mm:802), std::shuffle from Apple's libc++ (mm:797) and
// std::uniform_real_distribution<double>(0,1) (mm:802-803).  mt19937 is fully
// specified by the C++ standard and is used as is.  The other two are library
// algorithms whose exact draw sequence must match libc++, so they are written
// out here (libstdc++'s std::shuffle consumes the stream differently).
#pragma once
#include <stdint.h>

#include <random>
#include <utility>

namespace axon {

// libc++ uniform_int_distribution<ptrdiff_t> over mt19937 for the range
// [0, rp-1], rp <= 2^32: w = ceil(log2(rp)) low bits of one 32-bit word per
// try, rejected while >= rp.
inline uint64_t libcxx_uniform_index(std::mt19937& g, uint64_t rp) {
    if (rp <= 1) return 0;
    uint32_t w = 64u - (uint32_t)__builtin_clzll(rp) - 1u;
    if ((rp & (~0ull >> (64u - w))) != 0) ++w;
    const uint32_t mask = (w >= 32) ? 0xFFFFFFFFu : (0xFFFFFFFFu >> (32u - w));
    uint64_t u;
    do { u = (uint64_t)((uint32_t)g() & mask); } while (u >= rp);
    return u;
}

// libc++ std::shuffle: for i = 0 .. n-2, swap a[i] with a[i + j],
// j uniform in [0, n-1-i] (no swap when j == 0).
inline void libcxx_shuffle(uint32_t* a, uint32_t n, std::mt19937& g) {
    if (n < 2) return;
    for (uint32_t i = 0; i + 1 < n; ++i) {
        const uint64_t j = libcxx_uniform_index(g, (uint64_t)(n - i));
        if (j != 0) std::swap(a[i], a[i + j]);
    }
}

// generate_canonical<double, 53> over mt19937: two words, the first is the
// low half; the sum is rounded to double and divided by 2^64.
inline double canonical01(std::mt19937& g) {
    double s = (double)(uint32_t)g();
    s += (double)(uint32_t)g() * 4294967296.0;
    return s / 18446744073709551616.0;
}

}  // namespace axon
