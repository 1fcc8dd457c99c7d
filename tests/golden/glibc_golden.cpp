// Writes golden vectors of the system C library's rand() (implicit srand(1)) and of
// libstdc++'s std::random_shuffle, the two routines that order the reference's table populate
// (patchtable.h:3148). Built and run by tests/golden/make_golden.py; its output is committed as
// tests/golden/glibc_rand.json.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

static uint64_t fnv1a(const std::vector<int> &v) {
    uint64_t h = 1469598103934665603ull;
    for (int x : v) {
        uint32_t u = (uint32_t)x;
        for (int b = 0; b < 4; b++) {
            h ^= (u >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    }
    return h;
}

int main() {
    // rand() stream of a fresh process (glibc seeds it with 1).
    const int N = 4200000;
    std::vector<int> r(N);
    for (int i = 0; i < N; i++) r[i] = rand();
    printf("{\n  \"rand_first\": [");
    for (int i = 0; i < 2000; i++) printf("%s%d", i ? ", " : "", r[i]);
    printf("],\n  \"rand_sparse_index\": [");
    for (int i = 0, k = 0; i < N; i += 99991, k++) printf("%s%d", k ? ", " : "", i);
    printf("],\n  \"rand_sparse_value\": [");
    for (int i = 0, k = 0; i < N; i += 99991, k++) printf("%s%d", k ? ", " : "", r[i]);
    printf("],\n  \"rand_fnv1a_4200000\": \"%llu\",\n", (unsigned long long)fnv1a(r));
    // random_shuffle(iota(n)), each from a freshly seeded stream.
    const int sizes[] = {10, 1000, 115149, 773937, 4165681};
    printf("  \"shuffle\": [\n");
    for (int s = 0; s < 5; s++) {
        int n = sizes[s];
        srand(1);
        std::vector<int> a(n);
        std::iota(a.begin(), a.end(), 0);
        std::random_shuffle(a.begin(), a.end());
        printf("    {\"n\": %d, \"fnv1a\": \"%llu\", \"head\": [", n, (unsigned long long)fnv1a(a));
        int m = n < 1000 ? n : 1000;
        for (int i = 0; i < m; i++) printf("%s%d", i ? ", " : "", a[i]);
        printf("], \"tail\": [");
        for (int i = n - m; i < n; i++) printf("%s%d", i > n - m ? ", " : "", a[i]);
        printf("]}%s\n", s < 4 ? "," : "");
    }
    printf("  ]\n}\n");
    return 0;
}
