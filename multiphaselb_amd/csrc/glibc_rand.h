// rand() of the GNU C library (random_r.c, TYPE_3: additive feedback generator with
// 31 words of state and separation 3), used to reproduce the draws of the reference's
// fixedRandomT boundary (fixedRandomT.c:37,81), which reseeds with srand(pid) on every
// call.  Written from the published description of the algorithm:
//   state[0] = seed (0 -> 1); state[i] = 16807 state[i-1] mod (2^31 - 1), i = 1..30
//   (Schrage's method); then state[f] += state[r] over a ring of 31 words with f
//   starting 3 ahead of r, output = new state[f] >> 1; the first 310 outputs are
//   discarded.  Tests compare it with the host C library's rand().
#pragma once

#include <cstdint>

namespace mplb {

inline void glibc_rand_fill(unsigned seed, int n, int32_t* out) {
    int32_t st[31];
    int64_t word = seed == 0 ? 1 : (int32_t)seed;
    st[0] = (int32_t)word;
    for (int i = 1; i < 31; i++) {
        const int64_t hi = word / 127773;
        const int64_t lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        st[i] = (int32_t)word;
    }
    int f = 3, r = 0;
    auto next = [&]() -> int32_t {
        const uint32_t v = (uint32_t)st[f] + (uint32_t)st[r];
        st[f] = (int32_t)v;
        f = (f + 1) % 31;
        r = (r + 1) % 31;
        return (int32_t)(v >> 1);
    };
    for (int i = 0; i < 310; i++) next();
    for (int i = 0; i < n; i++) out[i] = next();
}

}  // namespace mplb
