// The rand() stream that orders the reference's populate (std::random_shuffle with glibc
// rand(), never seeded, patchtable.h:3148). glibc's random(3) TYPE_3 generator: r[0] = seed,
// r[i] = 16807 r[i-1] mod (2^31-1) for i < 31, r[i] = r[i-31] for 31 <= i < 34, then
// r[i] = r[i-3] + r[i-31] mod 2^32; outputs start after 310 discarded values and are r >> 1.
// The table build always starts from seed 1 (the state of a fresh process), so a table's
// contents do not depend on earlier rand() calls in the caller's process.
//
// The recurrence is linear over Z/2^32 in the 31-word state (r[k-31], ..., r[k-1]), so the
// stream is generated in parallel: the host provides the state before the first output and the
// jump matrices A^(L*2^i) (A the one-step transition), and each GPU thread jumps to the start of
// its chunk of L outputs (populate.hip).
#include "pt_internal.h"

#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

// State (r[313], ..., r[343]) before the first output r[344].
void pt_glibc_initial_state(uint32_t seed, uint32_t state[31]) {
    uint32_t u[344];
    int32_t s[34];
    s[0] = (int32_t)(seed == 0 ? 1 : seed);
    for (int i = 1; i < 31; i++) {
        int64_t hi = s[i - 1] / 127773, lo = s[i - 1] % 127773;
        int64_t word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        s[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; i++) s[i] = s[i - 31];
    for (int i = 0; i < 34; i++) u[i] = (uint32_t)s[i];
    for (int i = 34; i < 344; i++) u[i] = u[i - 31] + u[i - 3];
    for (int i = 0; i < 31; i++) state[i] = u[313 + i];
}

namespace {

typedef std::vector<uint32_t> Mat;   // 31 x 31, row-major

Mat matmul(const Mat &A, const Mat &B) {
    Mat C(31 * 31, 0);
    for (int i = 0; i < 31; i++)
        for (int k = 0; k < 31; k++) {
            uint32_t a = A[i * 31 + k];
            if (!a) continue;
            for (int j = 0; j < 31; j++) C[i * 31 + j] += a * B[k * 31 + j];
        }
    return C;
}

}  // namespace

// Jump matrices A^(L * 2^i), i = 0..levels-1, as levels x 31 x 31 words (cached per L).
const uint32_t *pt_glibc_jump_matrices(int L, int levels) {
    static std::mutex mu;
    static int cached_L = 0, cached_levels = 0;
    static std::vector<uint32_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    if (cached_L == L && cached_levels >= levels) return cache.data();
    Mat A(31 * 31, 0);
    for (int i = 0; i < 30; i++) A[i * 31 + i + 1] = 1;   // shift
    A[30 * 31 + 0] = 1;                                    // r[k] = r[k-31] + r[k-3]
    A[30 * 31 + 28] = 1;
    Mat P(31 * 31, 0);   // A^L by repeated squaring
    for (int i = 0; i < 31; i++) P[i * 31 + i] = 1;
    Mat base = A;
    for (int e = L; e; e >>= 1) {
        if (e & 1) P = matmul(P, base);
        base = matmul(base, base);
    }
    cache.assign((size_t)levels * 31 * 31, 0);
    for (int lv = 0; lv < levels; lv++) {
        std::memcpy(&cache[(size_t)lv * 961], P.data(), 961 * 4);
        P = matmul(P, P);
    }
    cached_L = L;
    cached_levels = levels;
    return cache.data();
}
