// Host-side work-item arithmetic (no device code).
#pragma once

#include <algorithm>
#include <cstdint>

namespace hbt {

// Pairs (i, j) with i in [a, a + ni), j in [jb, j1) and i < j, in closed
// form: i < jb pairs with all of [jb, j1); i >= jb with (i, j1).
inline uint64_t tile_pairs(long long a, long long ni, long long jb, long long j1) {
    uint64_t cnt = 0;
    const long long ie = a + ni;
    const long long e1 = std::min(ie, jb);
    if (e1 > a && j1 > jb) cnt += (uint64_t)(e1 - a) * (uint64_t)(j1 - jb);
    const long long b2 = std::max(a, jb), e2 = std::min(ie, j1 - 1);
    if (e2 > b2) {  // i = b2 .. e2-1 contribute j1-1-i: the integers lo..hi
        const long long n = e2 - b2, hi = j1 - 1 - b2, lo = j1 - e2;
        cnt += (uint64_t)((lo + hi) * n / 2);
    }
    return cnt;
}

}  // namespace hbt
