// The rand() stream that orders the reference's populate (std::random_shuffle with glibc
// rand(), never seeded, patchtable.h:3148). glibc's random(3) TYPE_3 generator: r[0] = seed,
// r[i] = 16807 r[i-1] mod (2^31-1) for i < 31, r[i] = r[i-31] for 31 <= i < 34, then
// r[i] = r[i-3] + r[i-31] mod 2^32; outputs start after 310 discarded values and are r >> 1.
// The table build always starts from seed 1 (the state of a fresh process), so a table's
// contents do not depend on earlier rand() calls in the caller's process.
#include "pt_internal.h"

#include <cstdint>

void pt_glibc_rand_stream(uint32_t seed, int64_t n, int32_t *out) {
    uint32_t ring[34];
    {
        int32_t s[34];
        s[0] = (int32_t)(seed == 0 ? 1 : seed);
        for (int i = 1; i < 31; i++) {
            int64_t hi = s[i - 1] / 127773, lo = s[i - 1] % 127773;
            int64_t word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            s[i] = (int32_t)word;
        }
        for (int i = 31; i < 34; i++) s[i] = s[i - 31];
        for (int i = 0; i < 34; i++) ring[i] = (uint32_t)s[i];
    }
    // ring[k % 34] holds r[k]; advance through the 310 discarded values.
    int64_t k = 34;
    for (; k < 344; k++) ring[k % 34] = ring[(k - 31) % 34] + ring[(k - 3) % 34];
    // Unrolled by the ring length so every index is a constant offset.
    int64_t i = 0;
    int p = (int)(k % 34);
    while (i < n) {
        uint32_t v = ring[(p + 3) % 34] + ring[(p + 31) % 34];
        ring[p] = v;
        out[i++] = (int32_t)(v >> 1);
        p = p == 33 ? 0 : p + 1;
    }
}
