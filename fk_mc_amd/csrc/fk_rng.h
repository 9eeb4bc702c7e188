// fk_rng.h -- std::mt19937 and the three libstdc++ distributions the reference
// draws from, restated for device (and host) code so that a GPU chain consumes
// exactly the same random words, in the same order, as the reference chain.
//
// Reference draw sites:
//   move choice   std::discrete_distribution<>      src/mc_metropolis.cpp:41
//   site choice   std::uniform_int_distribution<>   src/moves.cpp:7,55
//   Metropolis u  std::uniform_real_distribution<>  src/mc_metropolis.cpp:43
// Generator: std::mt19937 seeded SEED+rank (src/mc_metropolis.cpp:25).
//
// Semantics of the libstdc++ (GCC 11) algorithms reproduced here:
//   * uniform_int over [a,b] with a 32-bit engine: Lemire's multiply-shift with
//     rejection of low words below (2^32 - range) % range.
//   * generate_canonical<double,53>: two engine words g1,g2;
//     u = (g1 + g2*2^32) / 2^64, clamped to nextafter(1,0).
//   * discrete_distribution with fewer than two weights draws nothing and
//     returns 0; otherwise one canonical draw and lower_bound on the
//     normalised cumulative weights (last entry forced to 1).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define FK_HD __host__ __device__ inline
#else
#define FK_HD inline
#endif

namespace fkrng {

constexpr int MT_N = 624;
constexpr int MT_M = 397;
constexpr int MT_WORDS = MT_N + 1;  // state words + position index

// state[0..623] = generator words, state[624] = position (0..624)
FK_HD void mt_seed(uint32_t* st, uint32_t seed) {
    st[0] = seed;
    for (int i = 1; i < MT_N; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    st[MT_N] = MT_N;
}

FK_HD void mt_twist(uint32_t* st) {
    for (int i = 0; i < MT_N; ++i) {
        uint32_t y = (st[i] & 0x80000000u) | (st[(i + 1) % MT_N] & 0x7fffffffu);
        uint32_t v = st[(i + MT_M) % MT_N] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        st[i] = v;
    }
    st[MT_N] = 0;
}

FK_HD uint32_t mt_next(uint32_t* st) {
    if (st[MT_N] >= (uint32_t)MT_N) mt_twist(st);
    uint32_t y = st[st[MT_N]];
    st[MT_N] += 1;
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// uniform_int_distribution<int>(a, b)
FK_HD int uniform_int(uint32_t* st, int a, int b) {
    uint32_t range = (uint32_t)(b - a) + 1u;
    uint64_t product = (uint64_t)mt_next(st) * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        uint32_t threshold = (uint32_t)(0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)mt_next(st) * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return a + (int)(uint32_t)(product >> 32);
}

// generate_canonical<double, 53> == uniform_real_distribution<double>(0,1)
FK_HD double canonical(uint32_t* st) {
    double g1 = (double)mt_next(st);
    double g2 = (double)mt_next(st);
    double sum = g1 + g2 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0.99999999999999988898;  // nextafter(1.0, 0.0)
    return r;
}

// discrete_distribution: cp = cumulative normalised weights (ncp entries,
// cp[ncp-1] == 1).  ncp < 2 means the distribution holds a single weight.
FK_HD int discrete(uint32_t* st, const double* cp, int ncp) {
    if (ncp < 2) return 0;
    double p = canonical(st);
    int lo = 0, hi = ncp;  // first index with cp[i] >= p
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cp[mid] < p) lo = mid + 1; else hi = mid;
    }
    return lo;
}

}  // namespace fkrng
