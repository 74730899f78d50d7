// Host-side sort of the upload path (no device code).
#pragma once

#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace hbt {

// Stable LSD radix sort of (key, index) pairs by the key's numerical order,
// 8 bits a pass, skipping the passes whose digit all keys share.  Keys are
// not NaN; -0.0 and +0.0 compare equal (ties keep the input order), so the
// result is that of a stable comparison sort by key.  a and b are scratch.
template <typename I>
inline void radix_sort_keys(std::vector<std::pair<double, I>> &v,
                            std::vector<std::pair<uint64_t, I>> &a,
                            std::vector<std::pair<uint64_t, I>> &b) {
    const size_t n = v.size();
    a.resize(n);
    b.resize(n);
    // IEEE order as unsigned order: flip every bit of a negative key, the
    // sign bit of a positive one
    for (size_t k = 0; k < n; k++) {
        const double x = v[k].first + 0.0;  // -0.0 -> +0.0
        uint64_t u;
        std::memcpy(&u, &x, 8);
        u = (u >> 63) ? ~u : (u | (1ull << 63));
        a[k] = {u, v[k].second};
    }
    for (int sh = 0; sh < 64 && n > 0; sh += 8) {
        size_t cnt[256] = {};
        for (size_t k = 0; k < n; k++) cnt[(a[k].first >> sh) & 255]++;
        if (cnt[(a[0].first >> sh) & 255] == n) continue;
        size_t pos[256], t = 0;
        for (int d = 0; d < 256; d++) {
            pos[d] = t;
            t += cnt[d];
        }
        for (size_t k = 0; k < n; k++) b[pos[(a[k].first >> sh) & 255]++] = a[k];
        a.swap(b);
    }
    for (size_t k = 0; k < n; k++) {
        uint64_t u = a[k].first;
        u = (u >> 63) ? (u & ~(1ull << 63)) : ~u;
        double x;
        std::memcpy(&x, &u, 8);
        v[k] = {x, a[k].second};
    }
}

}  // namespace hbt
