// ============================================================================================
// PatchTable ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's hot path (uva-graphics/patchtable), written from the
// reference's source to check the HIP implementation against. Only tests/, the smoke check
// and the CPU-baseline leg of bench.py load it; the product library never links or calls it.
//
// Every function cites the reference file:line it follows. Floating-point expressions keep the
// reference's operand order and precision (FP64 where the reference uses double literals), and
// the file is compiled with -ffp-contract=off so no fused multiply-add changes a rounding.
//
// Parity status per function (how each is pinned; see tests/test_oracle_*.py):
//   glibc rand / random_shuffle ... golden vectors from the system glibc/libstdc++ (tests/golden)
//   rgb2yuv ........................ closed form in numpy float64 -> float32
//   gck / gck recurrences .......... WH-basis known-answer test (gck.py:70-77), recurrence table
//                                    extracted from gck.h (tests/golden), brute-force 64-tap
//                                    correlation in float64 on small images
//   quantiser ...................... brute force (sort, quantiles, nearest centre), limit
//                                    invariants of scale_slices_to_limit
//   populate ....................... invariant: each cell holds its minimum-rank patch
//   set_default_nslices ............ closed form (patchtable.h:1806-1809) composed with the
//                                    pinned scale_slices_to_limit
//   dt prop (BFS) .................. hand-worked known answers on tiny tables; brute-force
//                                    nearest source on 400 random lattices
//   dt downsample .................. hand-worked known answers
//   dt euclidean ................... nearest source in 1-D; a hand-worked 2-D case of the
//                                    in-place index update (patchtable.h:4268)
//   patch SSD ...................... float64 recompute, SSE lane order restated
//   k-coherence .................... this library's own kNN (the reference uses FLANN, absent
//                                    and unpinned): parity unpinned vs the reference; pinned by
//                                    recall@10 >= 0.25 against exact descriptor-space kNN
//                                    (0.29 measured) and the list invariants (sorted, distinct,
//                                    min_dist, float64 SSD, masked patches padded)
//   lookup (raster) ................ invariants of stress.py:134-154; quality only
// ============================================================================================
#include "../include/patchtable_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::vector<float> fvec;
typedef std::vector<int32_t> ivec;

static const int TABLE_UNUSED = (1 << 30) - 1;   // patchtable.h:109-112
static const int NMAP = 16384;                    // PATCH_PARTITION_NMAP, patchtable.h:216
static const double DIST_INFINITY = 100000;       // patchtable.h:173
static const int PT_MFMA_BAND_ORC = 1;            // candidate rows above/below an anchor (mfma.hip)

static inline int XY_TO_INT(int x, int y) { return x | (y << 16); }   // patchtable.h:163
static inline int INT_TO_X(int v) { return v & 65535; }
static inline int INT_TO_Y(int v) { return v >> 16; }

// Host threads of the loops the reference runs under OpenMP (its -threads, default 1). Only the
// OpenMP build of the oracle (liboracle_omp.so, for the CPU baseline's all-core timing) runs
// more than one; the checker library is built without OpenMP.
static inline int orc_threads(const pt_params &p) {
#ifdef _OPENMP
    return p.threads > 1 ? p.threads : 1;
#else
    (void)p;
    return 1;
#endif
}
static inline int orc_thread_num() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

// --------------------------------------------------------------------------------------------
// glibc random(3), TYPE_3 additive feedback generator, as rand() after the implicit srand(1).
// Restates glibc stdlib/random_r.c (__srandom_r, __random_r): r[0]=seed, r[i]=16807*r[i-1]
// mod (2^31-1) for i<31, r[i]=r[i-31] for i in 31..33, then r[i]=r[i-3]+r[i-31] (mod 2^32);
// the first 310 values after the seed block are discarded and each output is r>>1.
// --------------------------------------------------------------------------------------------
struct GlibcRand {
    uint32_t r[34];
    int pos;   // index (mod 34) of the next value to produce
    explicit GlibcRand(uint32_t seed = 1) {
        int32_t s[344];
        s[0] = (int32_t)(seed == 0 ? 1 : seed);
        for (int i = 1; i < 31; i++) {
            int64_t hi = s[i - 1] / 127773, lo = s[i - 1] % 127773;
            int64_t word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            s[i] = (int32_t)word;
        }
        for (int i = 31; i < 34; i++) s[i] = s[i - 31];
        uint32_t u[344];
        for (int i = 0; i < 34; i++) u[i] = (uint32_t)s[i];
        for (int i = 34; i < 344; i++) u[i] = u[i - 31] + u[i - 3];
        for (int i = 0; i < 34; i++) r[i] = u[310 + i];
        pos = 0;   // r holds u[310..343]; next value is u[344]
    }
    int32_t next() {
        // u[k] = u[k-31] + u[k-3]; ring of the last 34 values, r[pos] = u[k-34].
        uint32_t v = r[(pos + 3) % 34] + r[(pos + 31) % 34];
        r[pos] = v;
        pos = (pos + 1) % 34;
        return (int32_t)(v >> 1);
    }
};

// libstdc++ std::random_shuffle(first, last) with rand() (bits/stl_algo.h:4568-4584):
// for i = 1..n-1: j = rand() % (i+1); if (i != j) swap(a[i], a[j]).
static void random_shuffle(int32_t *a, int n, GlibcRand &rng) {
    for (int i = 1; i < n; i++) {
        int j = rng.next() % (i + 1);
        if (i != j) std::swap(a[i], a[j]);
    }
}

// --------------------------------------------------------------------------------------------
// rgb2yuv (solver/array.h:1116-1134): double-precision weights, float storage.
// --------------------------------------------------------------------------------------------
static void rgb2yuv(const float *in, int h, int w, float *out) {
    for (long i = 0; i < (long)h * w; i++) {
        float R = in[i * 3 + 0], G = in[i * 3 + 1], B = in[i * 3 + 2];
        out[i * 3 + 0] = (float)(0.299 * R + 0.587 * G + 0.114 * B);
        out[i * 3 + 1] = (float)(-0.14713 * R + -0.28886 * G + 0.436 * B);
        out[i * 3 + 2] = (float)(0.615 * R + -0.51499 * G + -0.10001 * B);
    }
}

// --------------------------------------------------------------------------------------------
// Walsh-Hadamard basis in Gray-code order and the snake ordering of 2-D kernels
// (gck.py:37-67). Each consecutive pair of kernels in snake order is linked by one recurrence
// along x or y with offset delta and sign (gck.py:389-400).
// --------------------------------------------------------------------------------------------
static void wh_basis_rec(std::vector<int> &L, int level, std::vector<int> &alpha, int w,
                         std::vector<int> &sign_level, std::vector<std::vector<int>> &vecs,
                         std::vector<std::vector<int>> &alphas) {
    std::vector<int> L1 = L, L2 = L;
    for (int x : L) L1.push_back(sign_level[level] * x);
    for (int x : L) L2.push_back(-sign_level[level] * x);
    std::vector<int> a1 = alpha, a2 = alpha;
    a1.push_back(sign_level[level]);
    a2.push_back(-sign_level[level]);
    sign_level[level] = -sign_level[level];
    if ((int)L1.size() == w) {
        vecs.push_back(L1);
        vecs.push_back(L2);
        alphas.push_back(a1);
        alphas.push_back(a2);
    } else {
        wh_basis_rec(L1, level + 1, a1, w, sign_level, vecs, alphas);
        wh_basis_rec(L2, level + 1, a2, w, sign_level, vecs, alphas);
    }
}

static void wh_basis(int w, std::vector<std::vector<int>> &vecs, std::vector<std::vector<int>> &alphas) {
    int k = 0;
    while ((1 << k) < w) k++;
    std::vector<int> sign_level(k, 1), L(1, 1), alpha;
    vecs.clear();
    alphas.clear();
    wh_basis_rec(L, 0, alpha, w, sign_level, vecs, alphas);
}

static std::vector<std::pair<int, int>> snake_order(int w) {
    std::vector<std::pair<int, int>> L;
    for (int r = 0; r < w; r++)
        for (int c = 0; c < w; c++) L.push_back({r, c});
    auto key = [w](const std::pair<int, int> &rc) {
        int r = rc.first, c = rc.second, m = std::max(r, c);
        return m + 0.5 + (r - c + w) * (1.0 / (4 * w)) * (m % 2 == 0 ? -1 : 1);
    };
    std::stable_sort(L.begin(), L.end(), [&](const std::pair<int, int> &a, const std::pair<int, int> &b) {
        return key(a) < key(b);
    });
    return L;
}

struct Recur { int axis; int delta; int sign; };   // axis 0: along x, 1: along y

static std::vector<Recur> gck_recurrences(int w) {
    std::vector<std::vector<int>> vecs, alphas;
    wh_basis(w, vecs, alphas);
    auto order = snake_order(w);
    std::vector<Recur> out;
    out.push_back({-1, 0, 0});   // kernel 0: box filter
    for (size_t i = 1; i < order.size(); i++) {
        bool horiz = order[i].second != order[i - 1].second;
        const std::vector<int> &a1 = horiz ? alphas[order[i - 1].second] : alphas[order[i - 1].first];
        const std::vector<int> &a2 = horiz ? alphas[order[i].second] : alphas[order[i].first];
        int idx = 0;
        while (a1[idx] == a2[idx]) idx++;
        out.push_back({horiz ? 0 : 1, 1 << idx, a2[idx]});
    }
    return out;
}

// --------------------------------------------------------------------------------------------
// GCK projection gck_8<float,float,n,nchroma> (gck.h:29-1950).
// in: h x w x 3 (YUV). out: (h+7) x (w+7) x (n+2*nchroma), HWC.
//   H pass (gck.h:54-165):  t(y,0)=in(y-3,0); t(y,x)=(t(y,x-1)+in(y-3,x))-in(y-3,x-8)
//   V pass (gck.h:166-216): o(0,x)=t(3,x);    o(y,x)=(o(y-1,x)+t(y+3,x))-t(y-5,x)
//   kernel k>=1 (gck.h:222-1946): k(y,x) = ((s*k(p-d)) + k-1(y,x)) + s*k-1(p-d), p-d the
//   position delta back along the recurrence axis; out-of-range reads are get_zero()
//   (solver/array.h:703-708), i.e. element (0,0,0) times zero.
// --------------------------------------------------------------------------------------------
static inline int gck_channel(int k, int nchroma) { return k < nchroma ? 3 * k : 3 * nchroma + (k - nchroma); }

static void gck(const float *in, int h, int w, int n, int nchroma, float *out) {
    const int H = h + 7, W = w + 7, C = n + 2 * nchroma;
    std::vector<float> temp((size_t)H * W * 3);
    const float in_zero = in[0] * 0.0f;
    auto inz = [&](int y, int x, int c) -> float {
        return (y >= 0 && y < h && x >= 0 && x < w) ? in[((size_t)y * w + x) * 3 + c] : in_zero;
    };
    for (int y = 0; y < H; y++) {
        int ys = y - 3;
        for (int c = 0; c < 3; c++) temp[((size_t)y * W) * 3 + c] = inz(ys, 0, c);
        for (int x = 1; x < W; x++)
            for (int c = 0; c < 3; c++) {
                float prev = temp[((size_t)y * W + x - 1) * 3 + c];
                temp[((size_t)y * W + x) * 3 + c] = (prev + inz(ys, x, c)) - inz(ys, x - 8, c);
            }
    }
    const float t_zero = temp[0] * 0.0f;
    auto tz = [&](int y, int x, int c) -> float {
        return (y >= 0 && y < H) ? temp[((size_t)y * W + x) * 3 + c] : t_zero;
    };
    for (int x = 0; x < W; x++)
        for (int c = 0; c < 3; c++) out[((size_t)x) * C + c] = temp[((size_t)3 * W + x) * 3 + c];
    for (int y = 1; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int c = 0; c < 3; c++) {
                float prev = out[((size_t)(y - 1) * W + x) * C + c];
                out[((size_t)y * W + x) * C + c] = (prev + tz(y + 3, x, c)) - tz(y - 5, x, c);
            }
    std::vector<Recur> rec = gck_recurrences(8);
    const float o_zero = out[0] * 0.0f;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int k = 1; k < n; k++) {
                int cc = k < nchroma ? 3 : 1;
                int chk = gck_channel(k, nchroma), chp = gck_channel(k - 1, nchroma);
                const Recur &r = rec[k];
                int yy = r.axis == 1 ? y - r.delta : y, xx = r.axis == 0 ? x - r.delta : x;
                bool ok = yy >= 0 && xx >= 0;
                for (int c = 0; c < cc; c++) {
                    float a1 = ok ? out[((size_t)yy * W + xx) * C + chk + c] : o_zero;
                    float a2 = ok ? out[((size_t)yy * W + xx) * C + chp + c] : o_zero;
                    float cur = out[((size_t)y * W + x) * C + chp + c];
                    float s1 = r.sign < 0 ? -a1 : a1, s2 = r.sign < 0 ? -a2 : a2;
                    out[((size_t)y * W + x) * C + chk + c] = (s1 + cur) + s2;
                }
            }
}

// --------------------------------------------------------------------------------------------
// Exact 8x8 RGB SSD with the SSE lane order of patch_dist_exact (patchtable.h:330-373):
// four lane accumulators; per row, six 4-wide chunks are added in order; lanes summed last.
// --------------------------------------------------------------------------------------------
static inline float patch_ssd(const float *a, int aw, int ax, int ay, const float *b, int bw, int bx, int by) {
    float L[4] = {0.f, 0.f, 0.f, 0.f};
    for (int dy = 0; dy < 8; dy++) {
        const float *ar = a + ((size_t)(ay + dy) * aw + ax) * 3;
        const float *br = b + ((size_t)(by + dy) * bw + bx) * 3;
        for (int j = 0; j < 4; j++) {
            float acc = L[j];
            for (int n = 0; n < 6; n++) {
                float d = ar[4 * n + j] - br[4 * n + j];
                acc = acc + d * d;
            }
            L[j] = acc;
        }
    }
    return ((L[0] + L[1]) + L[2]) + L[3];
}

// --------------------------------------------------------------------------------------------
// Descriptor distance of descriptor mode, patch_dist_approx (patchtable.h:376-437) with
// TABLE_SSE: the first four dimensions as one 4-lane (a-b)^2; for C = 8, 12, 16, 20 the other
// 4-wide chunks are added lane-wise and the lanes summed last, otherwise the lanes are summed
// and the remaining dimensions added one by one. C >= 4 (the SSE path reads four values).
// --------------------------------------------------------------------------------------------
static inline float desc_dist(const float *a, const float *b, int C) {
    float L[4];
    for (int j = 0; j < 4; j++) {
        float e = a[j] - b[j];
        L[j] = e * e;
    }
    if (C == 8 || C == 12 || C == 16 || C == 20) {
        for (int n = 1; n < C / 4; n++)
            for (int j = 0; j < 4; j++) {
                float e = a[4 * n + j] - b[4 * n + j];
                L[j] = L[j] + e * e;
            }
        return ((L[0] + L[1]) + L[2]) + L[3];
    }
    float d = ((L[0] + L[1]) + L[2]) + L[3];
    for (int i = 4; i < C; i++) {
        float e = a[i] - b[i];
        d += e * e;
    }
    return d;
}

// --------------------------------------------------------------------------------------------
// Quantiser.
// --------------------------------------------------------------------------------------------
// int(float) as the x86 cvttss2si instruction computes it: out-of-range and NaN give INT_MIN.
static inline int x86_ftoi(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int)f;
}

struct Partition1D {                       // PatchPartition1D (patchtable.h:1131-1326)
    fvec centers, lsizes, rsizes;
    ivec to_bin;
    float min_val, max_val, range;
    int nslices;
    int get_bin(float v) const {           // patchtable.h:1298-1303
        int idx = x86_ftoi((v - min_val) * (float)(NMAP - 1) / range);
        if (idx < 0) idx = 0;
        else if (idx >= NMAP) idx = NMAP - 1;
        return to_bin[idx];
    }
};

// scale_slices_to_limit (patchtable.cpp:611-642) with product()/multiply() (util.cpp:186-206).
static std::vector<int> scale_slices_to_limit(std::vector<int> ns, double limit) {
    auto product = [](const std::vector<int> &L) {
        double a = 1.0;
        for (int v : L) a *= v;
        return a;
    };
    if (limit > 0) {
        std::vector<int> ns0(ns);
        double factor = 1.0;
        for (auto &v : ns) v = std::max(v, 1);
        while (product(ns) > limit) {
            factor *= 0.995;
            for (size_t i = 0; i < ns.size(); i++) ns[i] = int(ns0[i] * factor + 0.5);
            for (auto &v : ns) v = std::max(v, 1);
        }
        for (;;) {
            ns0 = ns;
            bool success = false;
            for (size_t i = 0; i < ns.size(); i++) {
                ns[i]++;
                if (product(ns) > limit) ns = ns0;
                else {
                    success = true;
                    ns0 = ns;
                }
            }
            if (!success) break;
        }
    }
    return ns;
}

// PatchPartition1D constructor, non-regular grid, randomize=false (patchtable.h:1172-1286).
// `vals` are the channel's values over the listed patches; only their sorted order matters.
static Partition1D make_partition(std::vector<float> vals, int nslices, float min_val, float max_val) {
    Partition1D P;
    P.nslices = nslices;
    P.min_val = min_val;
    P.max_val = max_val;
    P.range = max_val - min_val;
    std::sort(vals.begin(), vals.end());
    long n = (long)vals.size();
    P.centers.resize(nslices);
    P.lsizes.resize(nslices);
    P.rsizes.resize(nslices);
    for (int i = 0; i < nslices; i++) {
        long idx1 = (n - 1) * (long)i / nslices;
        long idx2 = (n - 1) * (long)(i + 1) / nslices;
        float cell_max = vals[idx2], cell_min = vals[idx1];
        P.centers[i] = (cell_min + cell_max) / 2;
    }
    for (int i = 0; i < nslices; i++) {
        float prev_center = (i - 1 >= 0) ? P.centers[i - 1] : min_val;
        float next_center = (i + 1 < nslices) ? P.centers[i + 1] : max_val;
        P.lsizes[i] = P.centers[i] - prev_center;
        if (i - 1 >= 0) P.lsizes[i] /= 2;
        P.rsizes[i] = next_center - P.centers[i];
        if (i + 1 < nslices) P.rsizes[i] /= 2;
    }
    P.to_bin.resize(NMAP);
    int icenter = 0;
    for (int i = 0; i < NMAP; i++) {
        double coord = min_val + (i + 0.5) * P.range / (NMAP);
        while (coord > P.centers[icenter]) {
            if (icenter + 1 < nslices) icenter++;
            else break;
        }
        int prev_center = icenter - 1;
        if (prev_center < 0) prev_center = 0;
        int next_center = icenter;
        if (next_center == 0) next_center++;
        if (nslices == 1) next_center = prev_center = 0;
        double mid = (P.centers[prev_center] + P.centers[next_center]) / 2.0;
        P.to_bin[i] = (coord < mid) ? prev_center : next_center;
    }
    return P;
}

// --------------------------------------------------------------------------------------------
// The built table (PatchTableFixedN members, patchtable.h:1484-1540).
// --------------------------------------------------------------------------------------------
struct Table {
    pt_params p;
    int h = 0, w = 0;                 // "image" size: b in image mode; descriptor size + 7 otherwise
    int ndims = 0, G = 0;             // descriptor dims, grid dims
    bool desc = false;                // descriptor mode (is_descriptor, patchtable.h:4398-4401)
    fvec b0;                          // RGB copy of b (patchtable.h:2628); zeros in descriptor mode
    fvec wh0;                         // (h+7) x (w+7) x ndims
    ivec allowed;                     // (h-7) x (w-7) allowed_patches, empty without a mask
    int allowed_index = 1;
    std::vector<int> nslices;
    std::vector<int64_t> stride;      // row-major over nslices, dim 0 slowest
    int64_t ncells = 0;
    std::vector<Partition1D> part;
    ivec table_pop;                   // after populate
    ivec table;                       // after DT
    int table_add_count = 0;
    long rand_used = 0;               // rand() outputs consumed by the populate shuffles
    int K = 0;                        // k-coherence list length
    ivec kpos;                        // (h-7) x (w-7) x K
    fvec kdist;
    std::vector<Table *> extra;       // the ntables-1 further tables (patchtable.h:3359-3372)
    bool keyed = false;               // learned-key table: wh0 is the key; extra[0] is the GCK fallback
    Table() {}
    Table(const Table &) = delete;
    ~Table() {
        for (Table *t : extra) delete t;
    }
    int bew() const { return w - 7; }
    int beh() const { return h - 7; }
    const float *bdesc(int xs, int ys) const { return &wh0[((size_t)(ys + 7) * (w + 7) + xs + 7) * ndims]; }
    bool is_allowed(int xs, int ys, int index) const {
        return allowed.empty() || allowed[(size_t)ys * bew() + xs] == index;
    }
    int cell_of(const float *desc) const {
        int64_t idx = 0;
        for (int i = 0; i < G; i++) idx += (int64_t)part[i].get_bin(desc[i]) * stride[i];
        return (int)idx;
    }
};

// Descriptor-mode input placed as the padded descriptor (get_padded_descriptor,
// patchtable.h:2498-2507): unpadded h_in x w_in x C goes to rows/columns 7.. of an
// (h_in+14) x (w_in+14) array (zeros elsewhere); a padded input is taken as is.
static void pad_descriptor(const pt_params &p, const float *in, int h_in, int w_in, int C, fvec &wh, int &h, int &w) {
    if (p.descriptor_padded) {
        h = h_in - 7;
        w = w_in - 7;
        wh.assign(in, in + (size_t)h_in * w_in * C);
        return;
    }
    h = h_in + 7;
    w = w_in + 7;
    wh.assign((size_t)(h + 7) * (w + 7) * C, 0.f);
    for (int y = 0; y < h_in; y++)
        std::copy(in + (size_t)y * w_in * C, in + (size_t)(y + 1) * w_in * C, &wh[((size_t)(y + 7) * (w + 7) + 7) * C]);
}

// Key of the database of a learned-key table, (h+7) x (w+7) x 8 (computed on the GPU and handed
// in by the caller: the table's construction from the key is what is checked).
static const float *g_key_hwc = nullptr;

// Descriptor of an RGB image: colour conversion then GCK (get_padded_descriptor,
// patchtable.h:2492-2546, and reduce_dim, :1687-1728, for dim_algo WH without filter_dims).
static void descriptor(const pt_params &p, const float *img, int h, int w, fvec &wh) {
    fvec yuv((size_t)h * w * 3);
    if (p.convert_colorspace) rgb2yuv(img, h, w, yuv.data());
    else std::copy(img, img + (size_t)h * w * 3, yuv.begin());
    wh.assign((size_t)(h + 7) * (w + 7) * p.ndims, 0.f);
    gck(yuv.data(), h, w, p.ndims - p.nchroma * 2, p.nchroma, wh.data());
}

// ---------------------------- distance transforms -------------------------------------------

// do_dt_prop (patchtable.h:2081-2165): level-synchronous multi-source BFS over the cell
// lattice with PropEdge labels (patchtable.h:1447-1458); first claim in list order wins.
static void dt_prop_bfs(Table &T, int dt_iters) {
    ivec &tab = T.table;
    int G = T.G;
    int64_t nel = T.ncells;
    struct Elem { int64_t gi; int32_t val; uint8_t label; };
    std::vector<Elem> cur, nxt;
    for (int64_t j = 0; j < nel; j++)
        if (tab[j] != TABLE_UNUSED) cur.push_back({j, tab[j], 0});
    int nlabels = 1 + G * 2;
    struct Edge { int dim, dir, label; };
    std::vector<std::vector<Edge>> next_label;
    auto mk = [&](int dim, int dir) { return Edge{dim, dir, 2 * dim + 1 + (dir + 1) / 2}; };
    std::vector<Edge> e0;
    for (int i = 0; i < G; i++) {
        e0.push_back(mk(i, -1));
        e0.push_back(mk(i, 1));
    }
    next_label.push_back(e0);
    for (int i = 0; i < G; i++)
        for (int dir = -1; dir <= 1; dir += 2) {
            std::vector<Edge> e;
            e.push_back(mk(i, dir));
            for (int j = i + 1; j < G; j++) {
                e.push_back(mk(j, -1));
                e.push_back(mk(j, 1));
            }
            next_label.push_back(e);
        }
    (void)nlabels;
    for (int iter = 0;; iter++) {
        if (dt_iters > 0 && iter > dt_iters) break;
        if (cur.empty()) break;
        nxt.clear();
        for (const Elem &el : cur) {
            for (const Edge &e : next_label[el.label]) {
                int64_t gp = el.gi + e.dir * T.stride[e.dim];
                if (gp >= 0 && gp < nel && tab[gp] == TABLE_UNUSED) {
                    int64_t dim_index = (el.gi / T.stride[e.dim]) % T.nslices[e.dim] + e.dir;
                    if (dim_index >= 0 && dim_index < T.nslices[e.dim]) {
                        tab[gp] = el.val;
                        nxt.push_back({gp, el.val, (uint8_t)e.label});
                    }
                }
            }
        }
        std::swap(cur, nxt);
    }
}

// Separable form of the same transform (the GPU algorithm): nearest filled cell in L1 over the
// cell lattice, ties to the smallest source cell index; optionally limited to distance
// max_dist (-1: none). Used by the tests to show it equals dt_prop_bfs.
static void dt_prop_separable(Table &T, int max_dist) {
    int64_t nel = T.ncells;
    const uint64_t INF = ~0ull;
    std::vector<uint64_t> key(nel);
    for (int64_t j = 0; j < nel; j++) key[j] = T.table[j] != TABLE_UNUSED ? (uint64_t)j : INF;
    for (int d = 0; d < T.G; d++) {
        int64_t st = T.stride[d], n = T.nslices[d];
        for (int64_t base = 0; base < nel; base++) {
            if ((base / st) % n != 0) continue;   // line start along dim d
            uint64_t run = INF;
            for (int64_t i = 0; i < n; i++) {
                uint64_t &k = key[base + i * st];
                if (run != INF) run += (1ull << 32);
                if (k < run) run = k;
                k = run;
            }
            run = INF;
            for (int64_t i = n - 1; i >= 0; i--) {
                uint64_t &k = key[base + i * st];
                if (run != INF) run += (1ull << 32);
                if (k < run) run = k;
                k = run;
            }
        }
    }
    ivec out(nel);
    for (int64_t j = 0; j < nel; j++) {
        uint64_t k = key[j];
        if (k == INF || (max_dist >= 0 && (int64_t)(k >> 32) > max_dist)) out[j] = TABLE_UNUSED;
        else out[j] = T.table[(int64_t)(k & 0xffffffffu)];
    }
    T.table.swap(out);
}

// do_dt_downsample (patchtable.h:2339-2361) with dt_downsample / dt_upsample
// (dt_downsample.h, generated for 3..N dims): halve every dimension ((s+1)/2) until all are 1;
// a coarse cell takes the first filled child in nested-loop order (dim 0 outermost); then going
// back up, each empty fine cell copies its parent.
static void dt_downsample(Table &T) {
    int G = T.G;
    std::vector<std::vector<int>> sizes;
    std::vector<ivec> pyr;
    sizes.push_back(T.nslices);
    pyr.push_back(T.table);
    auto strides_of = [&](const std::vector<int> &s) {
        std::vector<int64_t> st(G);
        int64_t acc = 1;
        for (int i = G - 1; i >= 0; i--) {
            st[i] = acc;
            acc *= s[i];
        }
        return st;
    };
    auto all_one = [&](const std::vector<int> &s) {
        for (int v : s) if (v != 1) return false;
        return true;
    };
    while (!all_one(sizes.back())) {
        const std::vector<int> &hs = sizes.back();
        std::vector<int> ls(G);
        for (int i = 0; i < G; i++) ls[i] = (hs[i] + 1) / 2;
        auto hst = strides_of(hs), lst = strides_of(ls);
        int64_t ln = 1;
        for (int v : ls) ln *= v;
        ivec lo(ln);
        const ivec &hi = pyr.back();
        std::vector<int> lc(G), hc(G);
        for (int64_t li = 0; li < ln; li++) {
            for (int i = 0; i < G; i++) lc[i] = (int)((li / lst[i]) % ls[i]);
            int nchild = 1;
            std::vector<int> lo_c(G), cnt(G);
            for (int i = 0; i < G; i++) {
                lo_c[i] = 2 * lc[i];
                cnt[i] = std::min(2 * (lc[i] + 1), hs[i]) - lo_c[i];
                nchild *= cnt[i];
            }
            int32_t v_out = TABLE_UNUSED;
            for (int c = 0; c < nchild; c++) {   // nested order, dim 0 outermost
                int rem = c;
                int64_t hidx = 0;
                for (int i = G - 1; i >= 0; i--) {
                    int ci = rem % cnt[i];
                    rem /= cnt[i];
                    hidx += (int64_t)(lo_c[i] + ci) * hst[i];
                }
                int32_t v = hi[hidx];
                if (!(v & (1 << 30)) && v != TABLE_UNUSED) {
                    v_out = v;
                    break;
                }
            }
            lo[li] = v_out;
        }
        sizes.push_back(ls);
        pyr.push_back(lo);
    }
    for (int lvl = (int)pyr.size() - 1; lvl >= 1; lvl--) {
        const std::vector<int> &hs = sizes[lvl - 1], &ls = sizes[lvl];
        auto hst = strides_of(hs), lst = strides_of(ls);
        int64_t hn = 1;
        for (int v : hs) hn *= v;
        ivec &hi = pyr[lvl - 1];
        const ivec &lo = pyr[lvl];
        for (int64_t hidx = 0; hidx < hn; hidx++) {
            if (hi[hidx] != TABLE_UNUSED) continue;
            int64_t lidx = 0;
            for (int i = 0; i < G; i++) lidx += (int64_t)(((hidx / hst[i]) % hs[i]) / 2) * lst[i];
            hi[hidx] = lo[lidx];
        }
    }
    T.table.swap(pyr[0]);
}

// euclidean_dt + do_dt_euclidean (patchtable.h:4244-4352): Felzenszwalb-Huttenlocher lower
// envelope along dim 0, then 1, ..., carrying the table value; float arithmetic, INF = 1e20.
static void euclidean_line(float *f, int n, float *dist, int *v, float *z, int *idx) {
    const float INF = 1E20;
    int k = 0;
    v[0] = 0;
    z[0] = -INF;
    z[1] = +INF;
    for (int q = 1; q <= n - 1; q++) {
        float s = ((f[q] + q * q) - (f[v[k]] + v[k] * v[k])) / (2 * q - 2 * v[k]);
        while (s <= z[k]) {
            k--;
            s = ((f[q] + q * q) - (f[v[k]] + v[k] * v[k])) / (2 * q - 2 * v[k]);
        }
        k++;
        v[k] = q;
        z[k] = s;
        z[k + 1] = +INF;
    }
    k = 0;
    for (int q = 0; q <= n - 1; q++) {
        while (z[k + 1] < q) k++;
        dist[q] = (q - v[k]) * (q - v[k]) + f[v[k]];
        idx[q] = idx[v[k]];
    }
}

static void dt_euclidean(Table &T) {
    const float INF = 1E20;
    int64_t nel = T.ncells;
    int G = T.G;
    fvec td(nel);
    for (int64_t j = 0; j < nel; j++) td[j] = T.table[j] == TABLE_UNUSED ? INF : 0.0f;
    int maxdim = 0;
    for (int v : T.nslices) maxdim = std::max(maxdim, v);
    fvec f(maxdim), dist(maxdim), z(maxdim + 1);
    std::vector<int> v(maxdim), idx(maxdim);
    std::vector<int64_t> outer(G);
    for (int inner = 0; inner < G; inner++) {
        int64_t product = 1;
        for (int dim = G - 1; dim >= 0; dim--)
            if (dim != inner) {
                outer[dim] = product;
                product *= T.nslices[dim];
            }
        for (int64_t j = 0; j < product; j++) {
            int64_t ai = 0;
            for (int dim = 0; dim < G; dim++)
                if (dim != inner) ai += ((j / outer[dim]) % T.nslices[dim]) * T.stride[dim];
            int64_t st = T.stride[inner];
            int n = T.nslices[inner];
            for (int q = 0; q < n; q++) {
                f[q] = td[ai + q * st];
                idx[q] = T.table[ai + q * st];
            }
            euclidean_line(f.data(), n, dist.data(), v.data(), z.data(), idx.data());
            for (int q = 0; q < n; q++) {
                td[ai + q * st] = dist[q];
                T.table[ai + q * st] = idx[q];
            }
        }
    }
}

// ---------------------------- k-coherence ---------------------------------------------------
// This library's k-nearest-neighbour lists (replaces FLANN, patchtable.h:2653-2881). For every
// database patch p:
//   1. candidates: the patches within `window` positions of p in two sort orders of all patches,
//      by the Morton interleave of their table bins and by the Morton interleave of the bins
//      offset by half the slice count;
//   2. valid candidates are at pixel distance^2 >= kcoherence_min_dist^2 from p
//      (patchtable.h:2758-2768);
//   3. the K valid candidates nearest in descriptor L2 (all ndims, summed in index order;
//      ties to the smaller raster index) are kept;
//   4. they are re-scored by exact RGB SSD and sorted by (SSD, XY_TO_INT) as the reference
//      sorts its (dist, (index, valid)) triples (patchtable.h:2815-2835).
// Lists shorter than K (tiny images) repeat their last entry, or p itself when empty.
static int morton_bits(const Table &T) {
    int mx = 1;
    for (int i = 0; i < T.G; i++) mx = std::max(mx, T.nslices[i] + (T.nslices[i] >> 1));
    int B = 0;
    while ((1 << B) < mx) B++;
    return B;
}

static uint64_t morton_key(const Table &T, const int *bins, int shifted, int B) {
    int Gk = T.G;
    while (Gk * B > 63) Gk--;   // one bit of the 64 is left for the allowed flag of the sort
    uint64_t key = 0;
    for (int b = B - 1; b >= 0; b--)
        for (int d = 0; d < Gk; d++) {
            uint32_t v = (uint32_t)bins[d] + (shifted ? (uint32_t)(T.nslices[d] >> 1) : 0u);
            key = (key << 1) | ((v >> b) & 1u);
        }
    return key;
}

// This library's k-coherence lists (the reference searches a FLANN kd-tree, patchtable.h:2653-2881):
// over the allowed database patches (TABLE_FOR_ALLOWED_PATCHES_STEP, patchtable.h:2711), sorted
// by two Morton keys of their table bins, the candidates of p are the patches within W sorted
// positions in either order at pixel distance >= kcoherence_min_dist; the K nearest in
// descriptor L2 are re-scored with the lookup distance (exact RGB SSD, or patch_dist_approx in
// descriptor mode) and sorted by (distance, XY_TO_INT). Short lists repeat their last entry;
// patches outside the mask get K entries (p, +inf), which end any walk at once.
static void build_kcoherence(Table &T) {
    const int K = T.p.kcoherence, W = T.p.kcoh_window;
    const int bew = T.bew(), beh = T.beh(), nd = T.ndims, Wp = T.w + 7;
    const int md2 = T.p.kcoherence_min_dist * T.p.kcoherence_min_dist;
    const long N = (long)bew * beh;
    T.K = K;
    T.kpos.assign((size_t)N * K, 0);
    T.kdist.assign((size_t)N * K, 0.f);
    if (K <= 0 || N == 0) return;
    const int B = morton_bits(T);
    std::vector<uint64_t> keys[2];
    std::vector<int32_t> perm[2], pos[2];
    std::vector<int> bins(T.G);
    std::vector<char> ok(N);
    for (int o = 0; o < 2; o++) keys[o].resize(N);
    for (long i = 0; i < N; i++) {
        int px = (int)(i % bew), py = (int)(i / bew);
        ok[i] = T.is_allowed(px, py, T.allowed_index);
        const float *d = &T.wh0[((size_t)(py + 7) * Wp + px + 7) * nd];
        for (int g = 0; g < T.G; g++) bins[g] = T.part[g].get_bin(d[g]);
        keys[0][i] = morton_key(T, bins.data(), 0, B);
        keys[1][i] = morton_key(T, bins.data(), 1, B);
    }
    long Na = 0;
    for (int o = 0; o < 2; o++) {
        perm[o].clear();
        pos[o].assign(N, -1);
        for (long i = 0; i < N; i++)
            if (ok[i]) perm[o].push_back((int32_t)i);
        const std::vector<uint64_t> &ky = keys[o];
        std::sort(perm[o].begin(), perm[o].end(), [&](int32_t a, int32_t b) {
            return ky[a] != ky[b] ? ky[a] < ky[b] : a < b;
        });
        Na = (long)perm[o].size();
        for (long r = 0; r < Na; r++) pos[o][perm[o][r]] = (int32_t)r;
    }
    struct Cand { float d; int32_t q; };
    // patches are independent: with several threads (OpenMP build) they are split between them,
    // with the same lists as one thread
    const int nthr = orc_threads(T.p);
    (void)nthr;
#pragma omp parallel num_threads(nthr) if (nthr > 1)
    {
    std::vector<Cand> cands;
#pragma omp for schedule(dynamic, 256)
    for (long i = 0; i < N; i++) {
        int px = (int)(i % bew), py = (int)(i / bew);
        if (!ok[i]) {
            for (int k = 0; k < K; k++) {
                T.kpos[(size_t)i * K + k] = XY_TO_INT(px, py);
                T.kdist[(size_t)i * K + k] = INFINITY;
            }
            continue;
        }
        const float *dp = &T.wh0[((size_t)(py + 7) * Wp + px + 7) * nd];
        cands.clear();
        for (int o = 0; o < 2; o++) {
            long r = pos[o][i];
            for (long t = r - W; t <= r + W; t++) {
                if (t < 0 || t >= Na || t == r) continue;
                int32_t q = perm[o][t];
                int qx = q % bew, qy = q / bew;
                int dx = qx - px, dy = qy - py;
                if (dx * dx + dy * dy < md2) continue;
                const float *dq = &T.wh0[((size_t)(qy + 7) * Wp + qx + 7) * nd];
                float dd = 0.f;
                for (int c = 0; c < nd; c++) {
                    float e = dp[c] - dq[c];
                    dd = dd + e * e;
                }
                cands.push_back({dd, q});
            }
        }
        std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) {
            return a.d != b.d ? a.d < b.d : a.q < b.q;
        });
        cands.erase(std::unique(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.q == b.q; }),
                    cands.end());
        int m = std::min((int)cands.size(), K);
        struct Sel { float ssd; int32_t xy; };
        std::vector<Sel> sel(m);
        for (int k = 0; k < m; k++) {
            int qx = cands[k].q % bew, qy = cands[k].q / bew;
            sel[k].ssd = T.desc ? desc_dist(T.bdesc(px, py), T.bdesc(qx, qy), nd)
                                : patch_ssd(T.b0.data(), T.w, px, py, T.b0.data(), T.w, qx, qy);
            sel[k].xy = XY_TO_INT(qx, qy);
        }
        std::sort(sel.begin(), sel.end(), [](const Sel &a, const Sel &b) {
            return a.ssd != b.ssd ? a.ssd < b.ssd : a.xy < b.xy;
        });
        for (int k = 0; k < K; k++) {
            Sel s = m > 0 ? sel[std::min(k, m - 1)] : Sel{0.f, XY_TO_INT(px, py)};
            T.kpos[(size_t)i * K + k] = s.xy;
            T.kdist[(size_t)i * K + k] = s.ssd;
        }
    }
    }
}

// ---------------------------- table construction --------------------------------------------
// PatchTableFixedN constructor (patchtable.h:2584-3373): descriptor (colour conversion + GCK, or
// the padded input in descriptor mode), quantiser, fast populate (random order), DT or
// fill_zeros, k-coherence, then the ntables-1 further tables. Only allowed patches take part
// when a mask is given (TABLE_FOR_ALLOWED_PATCHES, patchtable.h:120-145). `rand_skip` is the
// number of rand() outputs consumed before this table's shuffle (the further tables continue
// the process-wide rand() stream after the earlier tables' shuffles).
static Table *build(const pt_params &p_in, const float *b, int h_in, int w_in, int c_in, const int32_t *allowed,
                    long rand_skip = 0) {
    Table *T = new Table();
    T->p = p_in;
    pt_params &p = T->p;
    T->desc = p.is_descriptor != 0;
    if (T->desc) {   // patchtable.h:4398-4401
        p.ndims = c_in;
        p.nchroma = 0;
    }
    if (p.grid_ndims < 0) p.grid_ndims = p.ndims;      // patchtable.h:2603
    T->ndims = p.ndims;
    T->G = p.grid_ndims;
    if (T->desc) {
        pad_descriptor(p, b, h_in, w_in, c_in, T->wh0, T->h, T->w);
        T->b0.assign((size_t)T->h * T->w * 3, 0.f);
    } else {
        T->h = h_in;
        T->w = w_in;
        T->b0.assign(b, b + (size_t)h_in * w_in * 3);
        if (g_key_hwc) T->wh0.assign(g_key_hwc, g_key_hwc + (size_t)(h_in + 7) * (w_in + 7) * T->ndims);
        else descriptor(p, b, h_in, w_in, T->wh0);
    }
    const int h = T->h, w = T->w;
    const int Wp = w + 7, nd = T->ndims, G = T->G;
    const int bew = T->bew(), beh = T->beh();
    if (allowed) {
        T->allowed.assign(allowed, allowed + (size_t)bew * beh);
        T->allowed_index = p.allowed_index;
    }
    auto ok = [&](int x, int y) { return T->is_allowed(x - 7, y - 7, T->allowed_index); };   // padded coords
    // set_min_max (patchtable.h:1758-1778)
    std::vector<float> mn(G, (float)1e20), mx(G, (float)-1e20);
    for (int y = 7; y < h; y++)
        for (int x = 7; x < w; x++) {
            if (!ok(x, y)) continue;
            for (int i = 0; i < G; i++) {
                float v = T->wh0[((size_t)y * Wp + x) * nd + i];
                mn[i] = mn[i] < v ? mn[i] : v;
                mx[i] = mx[i] > v ? mx[i] : v;
            }
        }
    std::vector<float> range(G);
    for (int i = 0; i < G; i++) range[i] = mx[i] - mn[i];
    // set_default_nslices (patchtable.h:1780-1810) or explicit -nslices, then the limit.
    std::vector<int> ns(G);
    if (p.nslices_count == 0) {
        float max_range = range[0];
        for (int i = 1; i < G; i++) max_range = std::max(max_range, range[i]);
        double scale = 200.0 / max_range;
        for (int i = 0; i < G; i++) ns[i] = max_range > 0 ? int(range[i] * scale) : 0;
    } else {
        for (int i = 0; i < G; i++) ns[i] = p.nslices[i];
    }
    T->nslices = scale_slices_to_limit(ns, p.limit);
    T->stride.assign(G, 0);
    int64_t acc = 1;
    for (int i = G - 1; i >= 0; i--) {
        T->stride[i] = acc;
        acc *= T->nslices[i];
    }
    T->ncells = acc;
    // PatchPartition over the allowed patches of the partition grid (patchtable.h:1355-1385).
    for (int i = 0; i < G; i++) {
        std::vector<float> vals;
        vals.reserve((size_t)bew * beh);
        for (int y = 7; y < h; y += p.partition_step)
            for (int x = 7; x < w; x += p.partition_step)
                if (ok(x, y)) vals.push_back(T->wh0[((size_t)y * Wp + x) * nd + i]);
        T->part.push_back(make_partition(vals, T->nslices[i], mn[i], mx[i]));
    }
    // Populate (patchtable.h:3121-3159), TABLE_POPULATE_FAST.
    T->table.assign(T->ncells, TABLE_UNUSED);
    std::vector<int32_t> order;
    for (int y = 7; y < h; y++)
        for (int x = 7; x < w; x++)
            if (ok(x, y)) order.push_back(XY_TO_INT(x, y));
    T->rand_used = rand_skip;
    if (p.populate_random) {
        GlibcRand rng(1);
        for (long k = 0; k < rand_skip; k++) rng.next();
        random_shuffle(order.data(), (int)order.size(), rng);
        if (order.size() > 1) T->rand_used += (long)order.size() - 1;
    }
    T->table_add_count = 0;
    for (int32_t pos : order) {
        int x = INT_TO_X(pos), y = INT_TO_Y(pos);
        int cell = T->cell_of(&T->wh0[((size_t)y * Wp + x) * nd]);
        if (T->table[cell] == TABLE_UNUSED) {
            T->table[cell] = pos;
            T->table_add_count++;
        }
    }
    T->table_pop = T->table;
    // DT dispatch (patchtable.h:3254-3295).
    bool zeros_filled = false;
    if (p.run_dt) {
        if (p.dt_algo == PT_DT_ALGO_PROP) {
            dt_prop_bfs(*T, p.dt_iters);
            zeros_filled = p.dt_iters <= 0;
        } else if (p.dt_algo == PT_DT_ALGO_DOWNSAMPLE) {
            dt_downsample(*T);
            zeros_filled = true;
        } else if (p.dt_algo == PT_DT_ALGO_EUCLIDEAN) {
            dt_euclidean(*T);
            zeros_filled = true;
        }
    }
    if (!zeros_filled) {   // fill_zeros (patchtable.h:2063-2079)
        int32_t vfill = XY_TO_INT(7, 7);
        for (auto &v : T->table)
            if (v == TABLE_UNUSED) v = vfill;
    }
    // the MFMA lookup (mfma_rounds > 0) walks no k-coherence lists, so none are built
    if (p.kcoherence > 0 && !(p.mfma_rounds > 0 && !T->desc)) build_kcoherence(*T);
    // Further tables (patchtable.h:3359-3372): limit / table_ratio^i, no k-coherence.
    long used = T->rand_used;
    for (int i = 1; i < p.ntables; i++) {
        pt_params pc = p;
        pc.limit = (double)(int)(pc.limit * (1.0 / pow(p.table_ratio, i)));
        pc.kcoherence = 0;
        pc.ntables = 1;
        Table *E = build(pc, b, h_in, w_in, c_in, allowed, used);
        used = E->rand_used;
        T->extra.push_back(E);
    }
    return T;
}

// ---------------------------- query side ---------------------------------------------------
// Descriptor of the query (get_padded_descriptor, patchtable.h:2492-2546): GCK of the colour-
// converted RGB image, or the (padded) input descriptor in descriptor mode.
struct QuerySide {
    int ah = 0, aw = 0;          // "image" size (descriptor size + 7 in descriptor mode)
    int C = 0;
    const float *a = nullptr;    // RGB query (image mode)
    fvec wh;                     // (ah+7) x (aw+7) x C
    fvec xwh;                    // learned-key tables: the GCK descriptor for the fallback table
    int xC = 0;
    int ann_w() const { return aw - 7; }
    int ann_h() const { return ah - 7; }
    const float *qdesc(int x, int y) const { return &wh[((size_t)(y + 7) * (aw + 7) + x + 7) * C]; }
    // descriptor indexing the further table E (its own for the fallback of a learned-key table)
    const float *xdesc(int x, int y) const {
        return xwh.empty() ? qdesc(x, y) : &xwh[((size_t)(y + 7) * (aw + 7) + x + 7) * xC];
    }
};

// Query key of a learned-key table, (h+7) x (w+7) x 8 (computed on the GPU, handed in).
static const float *g_query_key = nullptr;

static void query_side(const Table &T, const float *a, int h_in, int w_in, int c_in, QuerySide &Q) {
    Q.C = T.ndims;
    if (T.keyed) {
        Q.ah = h_in;
        Q.aw = w_in;
        Q.a = a;
        Q.wh.assign(g_query_key, g_query_key + (size_t)(h_in + 7) * (w_in + 7) * T.ndims);
        Q.xC = T.extra[0]->ndims;
        descriptor(T.extra[0]->p, a, h_in, w_in, Q.xwh);
    } else if (T.desc) {
        pad_descriptor(T.p, a, h_in, w_in, c_in, Q.wh, Q.ah, Q.aw);
    } else {
        Q.ah = h_in;
        Q.aw = w_in;
        Q.a = a;
        descriptor(T.p, a, h_in, w_in, Q.wh);
    }
}

// TABLE_LOOKUP_PATCH_DIST (patchtable.h:315-320): exact RGB SSD, or patch_dist_approx of the
// descriptors in descriptor mode.
static inline float lookup_dist(const Table &T, const QuerySide &Q, int x, int y, int xs, int ys) {
    return T.desc ? desc_dist(Q.qdesc(x, y), T.bdesc(xs, ys), Q.C)
                  : patch_ssd(Q.a, Q.aw, x, y, T.b0.data(), T.w, xs, ys);
}

static long rand_used_total(const Table &T) { return T.extra.empty() ? T.rand_used : T.extra.back()->rand_used; }

// ---------------------------- lookup (reference raster order) -------------------------------
// lookup_templated<is_descriptor, use_allowed_patches> (patchtable.h:3468-4217) with
// kcoherence_step 1 and prop_dist as given: ann_prev initialisation and temporal rescaling,
// table candidates of every table, k-coherence walk with the triangle test, spatial or random
// search, and the propagation scatter (summed-area table in image mode, patch_dist_approx in
// descriptor mode) with the border fill-in. rand() continues after the tables' shuffles.
static void lookup_raster(const Table &T, const float *a_in, int h_in, int w_in, int c_in, double *ann,
                          const double *ann_prev, const float *sv) {
    const pt_params &p = T.p;
    QuerySide Q;
    query_side(T, a_in, h_in, w_in, c_in, Q);
    const int ann_h = Q.ann_h(), ann_w = Q.ann_w();
    const int b0_ew = T.bew(), b0_eh = T.beh();
    const int a_w = Q.aw, a_h = Q.ah, b_w = T.w, b_h = T.h;
    const bool is_desc = T.desc, use_allowed = !T.allowed.empty();
    const float *a = Q.a;
    const float *b = T.b0.data();
    GlibcRand rng(1);
    for (long k = 0; k < rand_used_total(T); k++) rng.next();
    // With several threads the rows of a pass run concurrently, as under the reference's
    // `omp parallel for schedule(dynamic, 16)` (patchtable.h:3617-3618): a row reads matches that
    // other rows are writing, so the result depends on timing, and each thread draws from its
    // own copy of the generator. With one thread this is the reference's raster order exactly.
    const int nthr = orc_threads(p);
    std::vector<GlibcRand> trng(nthr, rng);
    auto allowed_at = [&](int xs, int ys) { return T.is_allowed(xs, ys, p.allowed_index); };
    auto dist = [&](int x, int y, int xs, int ys) { return lookup_dist(T, Q, x, y, xs, ys); };
    auto ANN = [&](int yy, int xx) -> double * { return &ann[((size_t)yy * ann_w + xx) * 3]; };
    const double prop_scale = 1.0 / (1 + p.coherence_spatial);
    const double temporal_scale = 1.0 / (1 + p.coherence_temporal);
    fvec dist_orig;
    if (!ann_prev) {
        for (int y = 0; y < ann_h; y++)
            for (int x = 0; x < ann_w; x++) {
                double *q = ANN(y, x);
                q[0] = b0_ew + 1;
                q[1] = 0;        // left uninitialised by the reference; any value behaves the same
                q[2] = DIST_INFINITY;
            }
    } else {   // patchtable.h:3539-3585
        if (p.recalc_dist_temporal) dist_orig.assign((size_t)ann_w * ann_h, 0.f);
        for (int y = 0; y < ann_h; y++)
            for (int x = 0; x < ann_w; x++) {
                const size_t i = (size_t)y * ann_w + x;
                const double *pp = &ann_prev[i * 3];
                int xsrc = (int)pp[0], ysrc = (int)pp[1];
                if (xsrc < 0) xsrc = 0;   // sanitize_input (the reference asserts in bounds otherwise)
                else if (xsrc >= b0_ew) xsrc = b0_ew - 1;
                if (ysrc < 0) ysrc = 0;
                else if (ysrc >= b0_eh) ysrc = b0_eh - 1;
                double *q = ANN(y, x);
                q[0] = xsrc;
                q[1] = ysrc;
                if (p.recalc_dist_temporal) {
                    float dc = dist(x, y, xsrc, ysrc);
                    double tsc = sv ? (double)sv[i] : temporal_scale;
                    dist_orig[i] = (float)(dc * tsc);
                    q[2] = dist_orig[i];
                } else {
                    q[2] = pp[2] * temporal_scale;
                }
            }
    }
    static const int dx_table[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
    static const int dy_table[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    const int qs = p.query_step;
    const int prop_w = (qs + p.prop_dist) * 2 + 1;
    const int sat_w = 8 + prop_w + 1 - 1;
    const int sat_w_nopad = sat_w - 1;
    const bool do_rs_or_spatial = p.do_rs || p.spatial;
    for (int prop_iter = 0; prop_iter < p.prop_iters; prop_iter++) {
        bool reverse = prop_iter % 2;
        bool do_lookup = prop_iter == 0 && p.do_table_lookup;
        bool do_kcoh = p.kcoherence > 0 && (p.kcoherence_iter < 0 || p.kcoherence_iter == prop_iter);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthr) if (nthr > 1)
        for (int y0 = 0; y0 < ann_h; y0 += qs) {
            int y = reverse ? (ann_h - 1 - y0) : y0;
            GlibcRand &rg = nthr > 1 ? trng[orc_thread_num()] : rng;
            std::vector<float> sat(sat_w * sat_w, 0.f);
            for (int x0 = 0; x0 < ann_w; x0 += qs) {
                int x = reverse ? (ann_w - 1 - x0) : x0;
                int xbest = (int)ANN(y, x)[0];
                int ybest = (int)ANN(y, x)[1];
                float dbest = (float)(ANN(y, x)[2] * prop_scale);
                auto consider = [&](int xs, int ys) {
                    float d = dist(x, y, xs, ys);
                    if (d < dbest) {
                        dbest = d;
                        xbest = xs;
                        ybest = ys;
                    }
                };
                if (do_lookup) {   // patchtable.h:3664-3737
                    const float *whp = Q.qdesc(x, y);
                    int v = T.table[T.cell_of(whp)];
                    int xsrc = INT_TO_X(v) - 7, ysrc = INT_TO_Y(v) - 7;
                    if (xsrc != xbest || ysrc != ybest) consider(xsrc, ysrc);
                    int xprev = xsrc, yprev = ysrc;
                    for (const Table *E : T.extra) {
                        int v1 = E->table[E->cell_of(Q.xdesc(x, y))];
                        int xs1 = INT_TO_X(v1) - 7, ys1 = INT_TO_Y(v1) - 7;
                        if ((xs1 != xbest || ys1 != ybest) && (xs1 != xprev || ys1 != yprev)) consider(xs1, ys1);
                        xprev = xs1;
                        yprev = ys1;
                    }
                }
                if (do_kcoh) {   // patchtable.h:3748-3851
                    const int32_t *kp = &T.kpos[((size_t)ybest * b0_ew + xbest) * T.K];
                    const float *kd = &T.kdist[((size_t)ybest * b0_ew + xbest) * T.K];
                    for (int k = 0; k < p.kcoherence; k++) {
                        if (kd[k] >= p.triangle_factor * dbest) break;
                        consider(INT_TO_X(kp[k]), INT_TO_Y(kp[k]));
                    }
                }
                if (do_rs_or_spatial) {   // patchtable.h:3853-3900
                    if (p.spatial) {
                        int symin = ybest - p.spatial, symax = ybest + p.spatial + 1;
                        int sxmin = xbest - p.spatial, sxmax = xbest + p.spatial + 1;
                        int xb0 = xbest, yb0 = ybest;
                        if (symin < 0) symin = 0;
                        else if (symax > b0_eh) symax = b0_eh;
                        if (sxmin < 0) sxmin = 0;
                        else if (sxmax > b0_ew) sxmax = b0_ew;
                        if (symax > b0_eh) symax = b0_eh;   // only differs where the reference reads out of bounds
                        if (sxmax > b0_ew) sxmax = b0_ew;
                        for (int ys = symin; ys < symax; ys++)
                            for (int xs = sxmin; xs < sxmax; xs++)
                                if ((xs != xb0 || ys != yb0) && allowed_at(xs, ys)) consider(xs, ys);
                    } else {
                        int r = rg.next() & 7;
                        int xs = xbest + dx_table[r], ys = ybest + dy_table[r];
                        if (xs >= 0 && xs < b0_ew && ys >= 0 && ys < b0_eh && allowed_at(xs, ys)) consider(xs, ys);
                    }
                }
                if (p.do_prop) {   // patchtable.h:3903-4148
                    int ax_shift = x - qs, ay_shift = y - qs;
                    int bx_shift = xbest - qs, by_shift = ybest - qs;
                    int dy_min = 0, dy_max = sat_w_nopad, dx_min = 0, dx_max = sat_w_nopad;
                    if (ax_shift + dx_min < 0) dx_min = -ax_shift;
                    else if (ax_shift + dx_max > a_w) dx_max = a_w - ax_shift;
                    if (ay_shift + dy_min < 0) dy_min = -ay_shift;
                    else if (ay_shift + dy_max > a_h) dy_max = a_h - ay_shift;
                    if (bx_shift + dx_min < 0) dx_min = -bx_shift;
                    else if (bx_shift + dx_max > b_w) dx_max = b_w - bx_shift;
                    if (by_shift + dy_min < 0) dy_min = -by_shift;
                    else if (by_shift + dy_max > b_h) dy_max = b_h - by_shift;
                    if (!is_desc) {
                        if (dx_min > 0 || dy_min > 0) {
                            if (dx_min > 0)
                                for (int dy = dy_min > 0 ? dy_min - 1 : dy_min; dy < dy_max; dy++) sat[(dy + 1) * sat_w + dx_min] = 0;
                            if (dy_min > 0)
                                for (int dx = dx_min; dx < dx_max; dx++) sat[(dy_min)*sat_w + dx + 1] = 0;
                        }
                        for (int dy = dy_min; dy < dy_max; dy++) {
                            float row_sum = 0;
                            for (int dx = dx_min; dx < dx_max; dx++) {
                                const float *ap = &a[((size_t)(ay_shift + dy) * a_w + ax_shift + dx) * 3];
                                const float *bp = &b[((size_t)(by_shift + dy) * b_w + bx_shift + dx) * 3];
                                float d0 = ap[0] - bp[0], d1 = ap[1] - bp[1], d2 = ap[2] - bp[2];
                                row_sum += d0 * d0 + d1 * d1 + d2 * d2;
                                sat[(dy + 1) * sat_w + 1 + dx] = row_sum + sat[dy * sat_w + 1 + dx];
                            }
                        }
                    }
                    int dy_max_patch = dy_max - 8 + 1, dx_max_patch = dx_max - 8 + 1;
                    bool missing = false;
                    for (int dy = dy_min; dy < dy_max_patch; dy++) {
                        const float *s1 = &sat[dy * sat_w], *s2 = &sat[(dy + 8) * sat_w];
                        int ysrc = dy + by_shift;
                        for (int dx = dx_min; dx < dx_max_patch; dx++) {
                            int xsrc = dx + bx_shift;
                            if (use_allowed && !allowed_at(xsrc, ysrc)) {
                                missing = true;
                                continue;
                            }
                            float dcur = is_desc ? desc_dist(Q.qdesc(ax_shift + dx, ay_shift + dy), T.bdesc(xsrc, ysrc), Q.C)
                                                 : s1[dx] - s1[dx + 8] - s2[dx] + s2[dx + 8];
                            double *ap = ANN(ay_shift + dy, ax_shift + dx);
                            if (dcur < ap[2]) {
                                ap[0] = xsrc;
                                ap[1] = ysrc;
                                ap[2] = dcur;
                            }
                        }
                    }
                    if (dx_max < sat_w_nopad || dy_max < sat_w_nopad || missing) {
                        int dy_start, dy_end, dx_start, dx_end;
                        if (missing) {
                            dx_start = 0; dx_end = qs * 2;
                            dy_start = 0; dy_end = qs * 2;
                            if (ax_shift + dx_start < 0) dx_start = -ax_shift;
                            if (ay_shift + dy_start < 0) dy_start = -ay_shift;
                            if (ax_shift + dx_end > ann_w) dx_end = ann_w - ax_shift;
                            if (ay_shift + dy_end > ann_h) dy_end = ann_h - ay_shift;
                        } else if (reverse) {
                            dx_start = 0; dx_end = qs;
                            dy_start = 0; dy_end = qs;
                            if (ax_shift + dx_start < 0) dx_start = -ax_shift;
                            if (ay_shift + dy_start < 0) dy_start = -ay_shift;
                        } else {
                            dx_start = qs; dx_end = qs * 2;
                            dy_start = qs; dy_end = qs * 2;
                            if (ax_shift + dx_end > ann_w) dx_end = ann_w - ax_shift;
                            if (ay_shift + dy_end > ann_h) dy_end = ann_h - ay_shift;
                        }
                        for (int dy = dy_start; dy < dy_end; dy++) {
                            int ysrc = ybest + (ay_shift + dy) - y;
                            if (ysrc < 0) ysrc = 0;
                            else if (ysrc >= b0_eh) ysrc = b0_eh - 1;
                            for (int dx = dx_start; dx < dx_end; dx++) {
                                double *ap = ANN(ay_shift + dy, ax_shift + dx);
                                if (ap[2] == DIST_INFINITY) {
                                    int xsrc = xbest + (ax_shift + dx) - x;
                                    if (xsrc < 0) xsrc = 0;
                                    else if (xsrc >= b0_ew) xsrc = b0_ew - 1;
                                    if (use_allowed)
                                        while (!allowed_at(xsrc, ysrc)) {
                                            xsrc = rg.next() % b0_ew;
                                            ysrc = rg.next() % b0_eh;
                                        }
                                    float d = dist(ax_shift + dx, ay_shift + dy, xsrc, ysrc);
                                    ap[2] = d;
                                    ap[0] = xsrc;
                                    ap[1] = ysrc;
                                }
                            }
                        }
                    }
                } else {
                    double *ap = ANN(y, x);
                    ap[0] = xbest;
                    ap[1] = ybest;
                    ap[2] = dbest;
                }
            }
        }
    }
    if (ann_prev && p.coherence_temporal != 0.0) {   // patchtable.h:4178-4193
        for (long i = 0; i < (long)ann_w * ann_h; i++) {
            double *q = &ann[i * 3];
            const double *pp = &ann_prev[i * 3];
            if (q[0] == pp[0] && q[1] == pp[1]) {
                if ((p.recalc_dist_temporal && q[2] == dist_orig[i]) ||
                    (!p.recalc_dist_temporal && q[2] == pp[2] * temporal_scale))
                    q[2] /= temporal_scale;
            }
        }
    }
}

// ---------------------------- lookup (this library's parallel form) ---------------------------
// Model of the GPU lookup. The reference visits the query patches of the lattice
// {x % query_step == 0, y % query_step == 0} in raster order and scatters each visited patch's
// offset to its (2*query_step+1)^2 neighbourhood. Here:
//   init: every patch starts from ann_prev (clamped; distance recomputed and scaled, or the
//     previous distance scaled, patchtable.h:3539-3585) or from nothing;
//   candidates (lattice patches, in parallel): the cell of every table (patchtable.h:3664-3737;
//     with no previous field the first table's candidate is taken unconditionally), then the
//     k-coherence walk of the best with the triangle test (patchtable.h:3779-3784), then one
//     random neighbour (do_rs without spatial);
//   rounds 1..R (lattice, double-buffered): for the 8 neighbours n = v + J*(dx,dy) of the
//     lattice the candidate NNF(n) - J*query_step*(dx,dy) if inside b and allowed; then the
//     k-coherence walk again when kcoherence_iter asks for it after pass 0 or when there are no
//     table candidates (do_table_lookup 0); then, from the second round on (in the candidate pass
//     when there are no rounds, in the first round when there is one), the spatial
//     window or a random neighbour. J runs 8, 4, 2, 1, then unit jumps;
//   gather (query_step > 1): every patch takes the best of the offsets of the lattice patches
//     within query_step of it (clamped into b, allowed only), starting from its initial match;
//   temporal: a patch that kept its initial match and distance gets it divided by the
//     temporal scale (patchtable.h:4178-4193).
// Candidates replace the best only when strictly closer, so NNF_DIST is the distance of the
// stored match. Skips that leave every result unchanged: a spatial window already searched
// around an unchanged best, a k-coherence list already walked at an unchanged best, and a
// neighbour's candidate that repeats the previous round's (same jump, neither changed).
static void jf_jumps(int rounds, std::vector<int> &J) {
    // e.g. rounds=6 -> 8,4,2,1,1,1
    J.clear();
    int big = rounds >= 4 ? 8 : (rounds == 3 ? 4 : (rounds == 2 ? 2 : 1));
    for (int j = big; j >= 1 && (int)J.size() < rounds; j >>= 1) J.push_back(j);
    while ((int)J.size() < rounds) J.push_back(1);
}

// Random neighbour of the random search (do_rs without spatial): a hash of the patch index and
// the round replaces the reference's rand() & 7 (patchtable.h:3887).
static inline int rs_pick(uint32_t i, uint32_t round) {
    uint32_t h = i * 0x9E3779B1u + round * 0x85EBCA77u + 0x165667B1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return (int)(h & 7u);
}

// Counters of distance evaluations in the parallel model (diagnostics): [0] tables +
// k-coherence, [1] spatial window of the candidate pass, [2 + 2r] neighbours (and k-coherence) of
// round r, [3 + 2r] spatial window of round r, [63] gather.
static long g_ssd_count[64];

// ---------------------------- lookup (this library's MFMA form) -----------------------------
// Model of the GPU MFMA lookup (mfma_rounds > 0; image mode, query_step 1, no mask). Matches are
// ranked by the integer SSD of the int8 images, q = round(255 clamp(v, 0, 1)) - 128 per value
// ("int8 SSD"); each patch's field holds its match and the match's score, and the exact FP32 SSD
// (the reference's lane order) of the final match is computed once at the end:
//   init: nothing, or the previous field's match (clamped) with score int8 SSD x w, w the
//     temporal weight (coherence_temporal_sv of the patch, else 1 / (1 + coherence_temporal));
//     its output distance is the reference's initial distance (the previous distance x w, or
//     with recalc_dist_temporal the exact SSD x w);
//   candidates: the cell of every table as lookup_parallel visits them (a cell equal to the
//     current match, or to the previous table's cell, is skipped), each taken when its int8 SSD
//     is strictly below the current score;
//   rounds r = 0..R-1 with jump J = mfma_jumps[r] (1 past the list's end) or, with no list,
//     J = max(1, 2^min(4, R-1) >> r), double-buffered, for every
//     group of 16 query patches x0 = 16k .. x0+15 of every row y:
//     sources (slots 0..63): the members (x0+i, y), the patches (x0+i, y-J) and (x0+i, y+J),
//       (x0+i-J, y) for i < 8 and (x0+i+J, y) for i >= 8, if inside the query; each proposes
//       its offset (match - position);
//     a round whose jump equals the previous round's skips every group none of whose valid
//       sources changed its match in the previous round (it would repeat that round's work and
//       keep every match; the skip only shows in the item counts);
//     anchors: the distinct offsets (count, first slot); up to mfma_anchors times the
//       uncovered one of largest count (ties: smallest first slot), which covers every offset
//       within 8 columns and PT_MFMA_BAND rows of it;
//     search: candidates (x0 + ox - 8 + j, y + oy + dy), j < 32, |dy| <= PT_MFMA_BAND, inside
//       b, in order o = (anchor * (2 PT_MFMA_BAND + 1) + dy + PT_MFMA_BAND) * 32 + j; each member
//       keeps the smallest (int8 SSD, o); when every score is at most the int8 SSD of its match
//       (no previous field, or no per-patch weights and coherence_temporal >= 0), a round after
//       the first does not search an anchor that the group searched in its last round with
//       anchors (the anchor keeps its index): those candidates cannot be taken, so only the
//       item counts change (pinned against the full search in the tests);
//     acceptance: a member whose best differs from its match takes it when its int8 SSD is
//       strictly below the match's score;
//   output: x, y and the exact FP32 SSD of the match; a patch still holding its initial
//     (previous-field) match keeps its initial distance, with the temporal rescaling of
//     lookup_parallel.
// The integer scores are exact, so this restates the matrix-core arithmetic bit for bit.
static inline int quant8(float v) { return (int)rintf(fminf(fmaxf(v, 0.f), 1.f) * 255.f) - 128; }

// (group, anchor) items of each round of the last lookup_mfma
static std::vector<long> g_mfma_items;
// 0: search every anchor (the full search the skip of searched anchors is checked against)
static int g_mfma_seen_skip = 1;

// jumps: pt_params.mfma_jumps (0-terminated; rounds past it use 1); empty: 16, 8, 4, 2, 1, then 1
static int mfma_jump_at(const int *jumps, int rounds, int r) {
    if (jumps[0] > 0) {
        for (int k = 0; k < PT_MFMA_MAX_JUMPS && jumps[k] > 0; k++)
            if (k == r) return jumps[k];
        return 1;
    }
    const int big = 1 << (rounds >= 5 ? 4 : (rounds >= 1 ? rounds - 1 : 0));
    const int j = big >> (r < 30 ? r : 30);
    return j > 1 ? j : 1;
}

static void lookup_mfma(const Table &T, const float *a_in, int h_in, int w_in, int c_in, double *ann,
                        const double *ann_prev, const float *sv) {
    const pt_params &p = T.p;
    const int R = PT_MFMA_BAND_ORC;
    QuerySide Q;
    query_side(T, a_in, h_in, w_in, c_in, Q);
    const int ah = Q.ann_h(), aw = Q.ann_w(), AW = Q.aw, AH = Q.ah;
    const int bew = T.bew(), beh = T.beh(), BW = T.w;
    const long N = (long)ah * aw;
    auto dist = [&](int x, int y, int xs, int ys) { return lookup_dist(T, Q, x, y, xs, ys); };
    const double temporal_scale = 1.0 / (1 + p.coherence_temporal);
    std::vector<int16_t> qa((size_t)AH * AW * 3), qb((size_t)T.h * BW * 3);
    for (size_t i = 0; i < qa.size(); i++) qa[i] = (int16_t)quant8(Q.a[i]);
    for (size_t i = 0; i < qb.size(); i++) qb[i] = (int16_t)quant8(T.b0[i]);
    auto ssd8 = [&](int x, int y, int xs, int ys) {
        int s = 0;
        for (int rr = 0; rr < 8; rr++) {
            const int16_t *qr = &qa[((size_t)(y + rr) * AW + x) * 3];
            const int16_t *br = &qb[((size_t)(ys + rr) * BW + xs) * 3];
            for (int e = 0; e < 24; e++) {
                const int dd = qr[e] - br[e];
                s += dd * dd;
            }
        }
        return s;
    };
    // init and candidates
    std::vector<int32_t> cx(N), cy(N), ix(N), iy(N);
    fvec cd(N), id(N);                     // scores; initial output distances
    std::vector<char> moved(N, 1);         // the match is not the initial (previous-field) one
    const bool have_prev = ann_prev != nullptr;
    for (long i = 0; i < N; i++) {
        const int x = (int)(i % aw), y = (int)(i / aw);
        bool have = false;
        int xb = 0, yb = 0;
        float db = 0.f;
        if (have_prev) {
            xb = (int)ann_prev[i * 3];
            yb = (int)ann_prev[i * 3 + 1];
            xb = xb < 0 ? 0 : (xb >= bew ? bew - 1 : xb);
            yb = yb < 0 ? 0 : (yb >= beh ? beh - 1 : yb);
            const double w = sv ? (double)sv[i] : temporal_scale;
            id[i] = p.recalc_dist_temporal ? (float)(dist(x, y, xb, yb) * w) : (float)(ann_prev[i * 3 + 2] * temporal_scale);
            db = (float)((double)ssd8(x, y, xb, yb) * w);
            have = true;
            moved[i] = 0;
        }
        ix[i] = xb;
        iy[i] = yb;
        if (p.do_table_lookup) {
            const int v = T.table[T.cell_of(Q.qdesc(x, y))];
            const int xs = INT_TO_X(v) - 7, ys = INT_TO_Y(v) - 7;
            if (!have || xs != xb || ys != yb) {
                const float d = (float)ssd8(x, y, xs, ys);
                if (!have || d < db) {
                    db = d;
                    xb = xs;
                    yb = ys;
                    moved[i] = 1;
                }
                have = true;
            }
            int xp = xs, yp = ys;
            for (const Table *E : T.extra) {
                const int v1 = E->table[E->cell_of(Q.xdesc(x, y))];
                const int x1 = INT_TO_X(v1) - 7, y1 = INT_TO_Y(v1) - 7;
                if ((x1 != xb || y1 != yb) && (x1 != xp || y1 != yp)) {
                    const float d = (float)ssd8(x, y, x1, y1);
                    if (d < db) {
                        db = d;
                        xb = x1;
                        yb = y1;
                        moved[i] = 1;
                    }
                }
                xp = x1;
                yp = y1;
            }
        }
        cx[i] = xb;
        cy[i] = yb;
        cd[i] = db;
    }
    const int rounds = p.mfma_rounds;
    std::vector<int32_t> nx(N), ny(N);
    fvec nd(N);
    g_mfma_items.assign(rounds, 0);
    std::vector<char> chg(N, 1);   // the previous round changed the match
    // each group's anchors of the last round in which it had any
    const int ngx = (aw + 15) / 16;
    std::vector<std::vector<std::pair<int, int>>> lastA((size_t)ngx * ah);
    const bool seen_ok = g_mfma_seen_skip && (!have_prev || (!sv && p.coherence_temporal >= 0.0));
    for (int r = 0; r < rounds; r++) {
        const int J = mfma_jump_at(p.mfma_jumps, rounds, r);
        const bool stable = r > 0 && mfma_jump_at(p.mfma_jumps, rounds, r - 1) == J;
        for (int y = 0; y < ah; y++)
            for (int x0 = 0; x0 < aw; x0 += 16) {
                int sox[64], soy[64];
                bool sval[64];
                bool any_moved = false;
                for (int sl = 0; sl < 64; sl++) {
                    int sx, sy;
                    if (sl < 16) sx = x0 + sl, sy = y;
                    else if (sl < 32) sx = x0 + sl - 16, sy = y - J;
                    else if (sl < 48) sx = x0 + sl - 32, sy = y + J;
                    else if (sl < 56) sx = x0 + (sl - 48) - J, sy = y;
                    else sx = x0 + (sl - 56) + 8 + J, sy = y;
                    sval[sl] = sx >= 0 && sy >= 0 && sx < aw && sy < ah;
                    sox[sl] = soy[sl] = 0;
                    if (sval[sl]) {
                        const long k = (long)sy * aw + sx;
                        sox[sl] = cx[k] - sx;
                        soy[sl] = cy[k] - sy;
                        any_moved = any_moved || chg[k];
                    }
                }
                if (stable && !any_moved)
                    for (int sl = 0; sl < 64; sl++) sval[sl] = false;
                // distinct offsets in order of first occurrence, with counts
                std::vector<int> dfirst, dcnt;
                for (int sl = 0; sl < 64; sl++) {
                    if (!sval[sl]) continue;
                    bool seen = false;
                    for (size_t k = 0; k < dfirst.size() && !seen; k++)
                        if (sox[dfirst[k]] == sox[sl] && soy[dfirst[k]] == soy[sl]) {
                            dcnt[k]++;
                            seen = true;
                        }
                    if (!seen) {
                        dfirst.push_back(sl);
                        dcnt.push_back(1);
                    }
                }
                std::vector<char> covered(dfirst.size(), 0);
                uint64_t best[16];
                for (int i = 0; i < 16; i++) best[i] = ~0ull;
                std::vector<int> aox, aoy;
                for (int a = 0; a < p.mfma_anchors; a++) {
                    int pick = -1;
                    for (size_t k = 0; k < dfirst.size(); k++) {
                        if (covered[k]) continue;
                        if (pick < 0 || dcnt[k] > dcnt[pick] || (dcnt[k] == dcnt[pick] && dfirst[k] < dfirst[pick]))
                            pick = (int)k;
                    }
                    if (pick < 0) break;
                    const int ox = sox[dfirst[pick]], oy = soy[dfirst[pick]];
                    aox.push_back(ox);
                    aoy.push_back(oy);
                    for (size_t k = 0; k < dfirst.size(); k++)
                        if (std::abs(sox[dfirst[k]] - ox) <= 8 && std::abs(soy[dfirst[k]] - oy) <= R) covered[k] = 1;
                    if (seen_ok && r > 0) {
                        bool seen = false;
                        for (const auto &q : lastA[(size_t)y * ngx + x0 / 16]) seen = seen || (q.first == ox && q.second == oy);
                        if (seen) continue;
                    }
                    g_mfma_items[r]++;
                    for (int dy = -R; dy <= R; dy++)
                        for (int j = 0; j < 32; j++) {
                            const int xs = x0 + ox - 8 + j, ys = y + oy + dy;
                            if (xs < 0 || ys < 0 || xs >= bew || ys >= beh) continue;
                            const uint32_t ord = (uint32_t)((a * (2 * R + 1) + dy + R) * 32 + j);
                            for (int i = 0; i < 16 && x0 + i < aw; i++) {
                                const uint64_t key = ((uint64_t)(uint32_t)ssd8(x0 + i, y, xs, ys) << 11) | ord;
                                if (key < best[i]) best[i] = key;
                            }
                        }
                }
                if (!aox.empty()) {
                    auto &la = lastA[(size_t)y * ngx + x0 / 16];
                    la.clear();
                    for (size_t q = 0; q < aox.size(); q++) la.push_back({aox[q], aoy[q]});
                }
                for (int i = 0; i < 16 && x0 + i < aw; i++) {
                    const long k = (long)y * aw + x0 + i;
                    int xb = cx[k], yb = cy[k];
                    float db = cd[k];
                    if (best[i] != ~0ull) {
                        const int ord = (int)(best[i] & 2047u), j = ord & 31, dyi = (ord >> 5) % (2 * R + 1),
                                  an = (ord >> 5) / (2 * R + 1);
                        const int xs = x0 + aox[an] - 8 + j, ys = y + aoy[an] - R + dyi;
                        const float s8 = (float)(uint32_t)(best[i] >> 11);
                        if ((xs != xb || ys != yb) && s8 < db) {
                            db = s8;
                            xb = xs;
                            yb = ys;
                            moved[k] = 1;
                        }
                    }
                    nx[k] = xb;
                    ny[k] = yb;
                    nd[k] = db;
                }
            }
        for (long k = 0; k < N; k++) chg[k] = nx[k] != cx[k] || ny[k] != cy[k];
        cx.swap(nx);
        cy.swap(ny);
        cd.swap(nd);
    }
    for (long i = 0; i < N; i++) {
        const int x = (int)(i % aw), y = (int)(i / aw);
        const float d0 = moved[i] ? dist(x, y, cx[i], cy[i]) : id[i];
        double d = d0;
        if (have_prev && p.coherence_temporal != 0.0 && (double)cx[i] == ann_prev[i * 3] &&
            (double)cy[i] == ann_prev[i * 3 + 1] && d0 == id[i])
            d = d / temporal_scale;
        ann[i * 3 + 0] = cx[i];
        ann[i * 3 + 1] = cy[i];
        ann[i * 3 + 2] = d;
    }
}

static void lookup_parallel(const Table &T, const float *a_in, int h_in, int w_in, int c_in, double *ann,
                            const double *ann_prev, const float *sv) {
    if (T.p.mfma_rounds > 0) {
        lookup_mfma(T, a_in, h_in, w_in, c_in, ann, ann_prev, sv);
        return;
    }
    for (auto &c : g_ssd_count) c = 0;
    const pt_params &p = T.p;
    QuerySide Q;
    query_side(T, a_in, h_in, w_in, c_in, Q);
    const int ann_h = Q.ann_h(), ann_w = Q.ann_w();
    const int bew = T.bew(), beh = T.beh();
    const int qs = p.query_step;
    const int vw = (ann_w + qs - 1) / qs, vh = (ann_h + qs - 1) / qs;
    const long N = (long)ann_h * ann_w, NV = (long)vw * vh;
    auto dist = [&](int x, int y, int xs, int ys) { return lookup_dist(T, Q, x, y, xs, ys); };
    auto allowed_at = [&](int xs, int ys) { return T.is_allowed(xs, ys, p.allowed_index); };
    const double temporal_scale = 1.0 / (1 + p.coherence_temporal);
    static const int DX[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
    static const int DY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    // initial field
    std::vector<int32_t> ix(N, 0), iy(N, 0);
    fvec id(N, 0.f);
    const bool have_prev = ann_prev != nullptr;
    if (have_prev)
        for (long i = 0; i < N; i++) {
            const int x = (int)(i % ann_w), y = (int)(i / ann_w);
            int xs = (int)ann_prev[i * 3], ys = (int)ann_prev[i * 3 + 1];
            xs = xs < 0 ? 0 : (xs >= bew ? bew - 1 : xs);
            ys = ys < 0 ? 0 : (ys >= beh ? beh - 1 : ys);
            ix[i] = xs;
            iy[i] = ys;
            if (p.recalc_dist_temporal) {
                const double tsc = sv ? (double)sv[i] : temporal_scale;
                id[i] = (float)(dist(x, y, xs, ys) * tsc);
            } else {
                id[i] = (float)(ann_prev[i * 3 + 2] * temporal_scale);
            }
        }
    const bool kcoh0 = p.kcoherence > 0 && (p.kcoherence_iter < 0 || p.kcoherence_iter == 0);
    // k-coherence walks in the rounds when kcoherence_iter asks for them after the first pass, and
    // without table candidates (the walk then has to follow the propagated matches).
    const bool kcohR = p.kcoherence > 0 && (p.kcoherence_iter < 0 || (p.kcoherence_iter >= 1 && p.kcoherence_iter < p.prop_iters) ||
                                            !p.do_table_lookup);
    const bool rs = !p.spatial && p.do_rs;
    // The spatial window runs from the second round on (in the candidate pass only when there
    // are no rounds, in the first round when there is one).
    const int nrounds = p.do_prop ? p.jf_rounds : 0;
    const bool spatial_cand = nrounds == 0;
    const int spatial_from = nrounds >= 2 ? 2 : 1;   // 1-based round index
    // the k-coherence walk of the best's list; returns true if it improved the best
    auto kwalk = [&](int x, int y, int &xb, int &yb, float &db, int slot) -> bool {
        const int32_t *kp = &T.kpos[((size_t)yb * bew + xb) * T.K];
        const float *kd = &T.kdist[((size_t)yb * bew + xb) * T.K];
        bool changed = false;
        for (int k = 0; k < p.kcoherence; k++) {
            if (kd[k] >= p.triangle_factor * db) break;
            int xs = INT_TO_X(kp[k]), ys = INT_TO_Y(kp[k]);
            float d = dist(x, y, xs, ys);
            g_ssd_count[slot]++;
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
                changed = true;
            }
        }
        return changed;
    };
    // candidates evaluated for the current patch in the current phase: evaluating one again
    // cannot change the result (rejected against a best that has only decreased since, or the
    // best itself), so the model skips repeats like the kernels do
    std::vector<int32_t> seen;
    auto is_seen = [&](int xs, int ys) { return std::find(seen.begin(), seen.end(), XY_TO_INT(xs, ys)) != seen.end(); };
    auto spatial = [&](int x, int y, int &xb, int &yb, float &db, int slot) -> bool {
        bool changed = false;
        int symin = yb - p.spatial, symax = yb + p.spatial + 1;
        int sxmin = xb - p.spatial, sxmax = xb + p.spatial + 1;
        int xb0 = xb, yb0 = yb;
        if (symin < 0) symin = 0;
        else if (symax > beh) symax = beh;
        if (sxmin < 0) sxmin = 0;
        else if (sxmax > bew) sxmax = bew;
        if (symax > beh) symax = beh;
        if (sxmax > bew) sxmax = bew;
        for (int ys = symin; ys < symax; ys++)
            for (int xs = sxmin; xs < sxmax; xs++)
                if ((xs != xb0 || ys != yb0) && allowed_at(xs, ys) && !is_seen(xs, ys)) {
                    float d = dist(x, y, xs, ys);
                    g_ssd_count[slot]++;
                    if (d < db) {
                        db = d;
                        xb = xs;
                        yb = ys;
                        changed = true;
                    }
                }
        return changed;
    };
    auto rsearch = [&](int x, int y, int &xb, int &yb, float &db, int round, int slot) -> bool {
        const int r = rs_pick((uint32_t)((long)y * ann_w + x), (uint32_t)round);
        const int xs = xb + DX[r], ys = yb + DY[r];
        if (xs < 0 || xs >= bew || ys < 0 || ys >= beh || !allowed_at(xs, ys)) return false;
        float d = dist(x, y, xs, ys);
        g_ssd_count[slot]++;
        if (d < db) {
            db = d;
            xb = xs;
            yb = ys;
            return true;
        }
        return false;
    };
    std::vector<int32_t> cx(NV), cy(NV), nx(NV), ny(NV);
    fvec cd(NV), ndist(NV);
    std::vector<uint8_t> cdone(NV), ndone(NV), cchg(NV, 1), nchg(NV), ckd(NV), nkd(NV);
    for (int vy = 0; vy < vh; vy++)
        for (int vx = 0; vx < vw; vx++) {
            const int x = vx * qs, y = vy * qs;
            const long i = (long)y * ann_w + x;
            bool have = have_prev;
            int xb = ix[i], yb = iy[i];
            float db = id[i];
            if (p.do_table_lookup) {
                const float *whp = Q.qdesc(x, y);
                int v = T.table[T.cell_of(whp)];
                int xs = INT_TO_X(v) - 7, ys = INT_TO_Y(v) - 7;
                if (!have || xs != xb || ys != yb) {
                    float d = dist(x, y, xs, ys);
                    g_ssd_count[0]++;
                    if (!have || d < db) {
                        db = d;
                        xb = xs;
                        yb = ys;
                    }
                    have = true;
                }
                int xp = xs, yp = ys;
                for (const Table *E : T.extra) {
                    int v1 = E->table[E->cell_of(Q.xdesc(x, y))];
                    int x1 = INT_TO_X(v1) - 7, y1 = INT_TO_Y(v1) - 7;
                    if ((x1 != xb || y1 != yb) && (x1 != xp || y1 != yp)) {
                        float d = dist(x, y, x1, y1);
                        g_ssd_count[0]++;
                        if (d < db) {
                            db = d;
                            xb = x1;
                            yb = y1;
                        }
                    }
                    xp = x1;
                    yp = y1;
                }
            }
            bool kd = false;
            if (kcoh0) kd = !kwalk(x, y, xb, yb, db, 0);
            bool done = false;
            seen.clear();
            if (p.spatial && spatial_cand) {
                const bool changed = spatial(x, y, xb, yb, db, 1);
                done = !changed;
                if (changed) kd = false;
            } else if (rs) {
                if (rsearch(x, y, xb, yb, db, 0, 1)) kd = false;
            }
            const long v = (long)vy * vw + vx;
            cx[v] = xb;
            cy[v] = yb;
            cd[v] = db;
            cdone[v] = done;
            ckd[v] = kd;
        }
    std::vector<int> J;
    jf_jumps(p.do_prop ? p.jf_rounds : 0, J);   // do_prop 0: no propagation
    for (size_t r = 0; r < J.size(); r++) {
        const int j = J[r];
        const bool same_jump = r > 0 && J[r - 1] == j;
        for (int vy = 0; vy < vh; vy++)
            for (int vx = 0; vx < vw; vx++) {
                const long v = (long)vy * vw + vx;
                const int x = vx * qs, y = vy * qs;
                int xb = cx[v], yb = cy[v];
                float db = cd[v];
                bool moved = false;
                seen.clear();
                for (int t = 0; t < 8; t++) {
                    const int qx = vx + j * DX[t], qy = vy + j * DY[t];
                    if (qx < 0 || qy < 0 || qx >= vw || qy >= vh) continue;
                    const long qv = (long)qy * vw + qx;
                    if (same_jump && !cchg[v] && !cchg[qv]) continue;
                    const int xs = cx[qv] - j * qs * DX[t], ys = cy[qv] - j * qs * DY[t];
                    if (xs < 0 || ys < 0 || xs >= bew || ys >= beh) continue;
                    if (xs == xb && ys == yb) continue;
                    if (!allowed_at(xs, ys)) continue;
                    if (is_seen(xs, ys)) continue;
                    seen.push_back(XY_TO_INT(xs, ys));
                    float d = dist(x, y, xs, ys);
                    g_ssd_count[2 + 2 * r]++;
                    if (d < db) {
                        db = d;
                        xb = xs;
                        yb = ys;
                        moved = true;
                    }
                }
                bool kd = ckd[v] && !moved;
                bool kchg = false;
                if (kcohR && !kd) {
                    kchg = kwalk(x, y, xb, yb, db, 2 + 2 * (int)r);
                    kd = !kchg;
                }
                bool done = cdone[v] && !moved && !kchg;
                bool schg = false;
                if (p.spatial && !done && (int)r + 1 >= spatial_from) {
                    schg = spatial(x, y, xb, yb, db, 3 + 2 * (int)r);
                    done = !schg;
                } else if (rs) {
                    schg = rsearch(x, y, xb, yb, db, (int)r + 1, 3 + 2 * (int)r);
                }
                if (schg) kd = false;
                nx[v] = xb;
                ny[v] = yb;
                ndist[v] = db;
                ndone[v] = done;
                nchg[v] = moved || kchg || schg;
                nkd[v] = kd;
            }
        cx.swap(nx);
        cy.swap(ny);
        cd.swap(ndist);
        cdone.swap(ndone);
        cchg.swap(nchg);
        ckd.swap(nkd);
    }
    // output: the lattice itself, or the gather over the lattice neighbours of every patch
    std::vector<int32_t> ox(N), oy(N);
    fvec od(N);
    if (qs == 1) {
        for (long i = 0; i < N; i++) {
            ox[i] = cx[i];
            oy[i] = cy[i];
            od[i] = cd[i];
        }
    } else {
        for (int y = 0; y < ann_h; y++)
            for (int x = 0; x < ann_w; x++) {
                const long i = (long)y * ann_w + x;
                bool have = have_prev;
                int xb = ix[i], yb = iy[i];
                float db = id[i];
                const int vy0 = y / qs - (y % qs == 0 ? 1 : 0), vy1 = y / qs + 1;
                const int vx0 = x / qs - (x % qs == 0 ? 1 : 0), vx1 = x / qs + 1;
                for (int vy = vy0 < 0 ? 0 : vy0; vy <= vy1 && vy < vh; vy++)
                    for (int vx = vx0 < 0 ? 0 : vx0; vx <= vx1 && vx < vw; vx++) {
                        const long v = (long)vy * vw + vx;
                        int xs = cx[v] + (x - vx * qs), ys = cy[v] + (y - vy * qs);
                        xs = xs < 0 ? 0 : (xs >= bew ? bew - 1 : xs);
                        ys = ys < 0 ? 0 : (ys >= beh ? beh - 1 : ys);
                        if (!allowed_at(xs, ys)) continue;
                        if (have && xs == xb && ys == yb) continue;
                        float d = dist(x, y, xs, ys);
                        g_ssd_count[63]++;
                        if (!have || d < db) {
                            db = d;
                            xb = xs;
                            yb = ys;
                            have = true;
                        }
                    }
                if (!have) {   // every proposal disallowed: the owner lattice patch's match
                    const long v = (long)(y / qs) * vw + x / qs;
                    xb = cx[v];
                    yb = cy[v];
                    db = dist(x, y, xb, yb);
                }
                ox[i] = xb;
                oy[i] = yb;
                od[i] = db;
            }
    }
    for (long i = 0; i < N; i++) {
        double d = od[i];
        if (have_prev && p.coherence_temporal != 0.0 && (double)ox[i] == ann_prev[i * 3] &&
            (double)oy[i] == ann_prev[i * 3 + 1] && od[i] == id[i])
            d = d / temporal_scale;
        ann[i * 3 + 0] = ox[i];
        ann[i * 3 + 1] = oy[i];
        ann[i * 3 + 2] = d;
    }
}

// ---------------------------- learned key (MLP 192 -> 64 -> 8) --------------------------------
// The key of this library's learned-key table: x = patch - 0.5 in bf16 (HWC order), h =
// ReLU(W1 x + b1), k = W2 h + b2 with W1, W2 rounded to bf16 and biases in float; sums here in
// double (the GPU accumulates in FP32 on MFMA / VALU, so the two agree to FP32 rounding).
static inline float bf16_round(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) u = (u | 0x400000u) & 0xffff0000u;
    else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static const int KEY_IN = 192, KEY_H = 64, KEY_OUT = 8;

static void key_patch(const float *img, int w, int x, int y, float *xv) {
    for (int dy = 0; dy < 8; dy++)
        for (int r = 0; r < 24; r++) xv[dy * 24 + r] = bf16_round(img[((size_t)(y + dy) * w + x) * 3 + r] - 0.5f);
}

// Weights are rounded to BF16 as the GPU stores them; a finite-difference check of the
// gradients switches the rounding off (the rounded loss is piecewise constant in the weights).
static bool g_key_round_weights = true;

struct KeyNet {
    std::vector<double> w1, b1, w2, b2;
    KeyNet(const float *W1, const float *B1, const float *W2, const float *B2)
        : w1(W1, W1 + KEY_H * KEY_IN), b1(B1, B1 + KEY_H), w2(W2, W2 + KEY_OUT * KEY_H), b2(B2, B2 + KEY_OUT) {
        if (g_key_round_weights) {
            for (auto &v : w1) v = bf16_round((float)v);
            for (auto &v : w2) v = bf16_round((float)v);
        }
    }
    // weights given in double (finite differences)
    KeyNet(const double *W1, const double *B1, const double *W2, const double *B2)
        : w1(W1, W1 + KEY_H * KEY_IN), b1(B1, B1 + KEY_H), w2(W2, W2 + KEY_OUT * KEY_H), b2(B2, B2 + KEY_OUT) {}
    // round_hidden: the ReLU'd hidden units enter layer 2 rounded to BF16 (nearest even), as the
    // GPU's inference forward feeds them to its BF16 layer-2 MFMA; training keeps them exact
    void forward(const float *x, double *z, double *k, bool round_hidden = false) const {
        for (int j = 0; j < KEY_H; j++) {
            double s = b1[j];
            for (int i = 0; i < KEY_IN; i++) s += w1[j * KEY_IN + i] * x[i];
            z[j] = s;
        }
        for (int o = 0; o < KEY_OUT; o++) {
            double s = b2[o];
            for (int j = 0; j < KEY_H; j++) {
                const double hv = z[j] > 0 ? z[j] : 0.0;
                s += w2[o * KEY_H + j] * (round_hidden ? (double)bf16_round((float)hv) : hv);
            }
            k[o] = s;
        }
    }
};

}  // namespace orc

// ============================================================================================
// C ABI of the oracle (loaded by tests with ctypes).
// ============================================================================================
using namespace orc;

extern "C" {

int orc_glibc_rand(uint32_t seed, int n, int32_t *out) {
    GlibcRand r(seed);
    for (int i = 0; i < n; i++) out[i] = r.next();
    return 0;
}

int orc_random_shuffle(int n, int32_t *a) {
    GlibcRand r(1);
    random_shuffle(a, n, r);
    return 0;
}

int orc_rgb2yuv(const float *in, int h, int w, float *out) {
    rgb2yuv(in, h, w, out);
    return 0;
}

int orc_wh_basis(int w, int *vecs_out, int *alphas_out) {
    std::vector<std::vector<int>> vecs, alphas;
    wh_basis(w, vecs, alphas);
    int k = (int)alphas[0].size();
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) vecs_out[i * w + j] = vecs[i][j];
    for (int i = 0; i < w; i++)
        for (int j = 0; j < k; j++) alphas_out[i * k + j] = alphas[i][j];
    return 0;
}

int orc_snake_order(int w, int *rows, int *cols) {
    auto L = snake_order(w);
    for (size_t i = 0; i < L.size(); i++) {
        rows[i] = L[i].first;
        cols[i] = L[i].second;
    }
    return 0;
}

int orc_gck_recurrences(int w, int *axis, int *delta, int *sign) {
    auto R = gck_recurrences(w);
    for (size_t i = 0; i < R.size(); i++) {
        axis[i] = R[i].axis;
        delta[i] = R[i].delta;
        sign[i] = R[i].sign;
    }
    return (int)R.size();
}

int orc_gck(const float *in, int h, int w, int n, int nchroma, float *out) {
    gck(in, h, w, n, nchroma, out);
    return 0;
}

int orc_descriptor(const pt_params *p, const float *img, int h, int w, float *out) {
    fvec wh;
    descriptor(*p, img, h, w, wh);
    std::copy(wh.begin(), wh.end(), out);
    return 0;
}

float orc_patch_ssd(const float *a, int aw, int ax, int ay, const float *b, int bw, int bx, int by) {
    return patch_ssd(a, aw, ax, ay, b, bw, bx, by);
}

int orc_scale_slices_to_limit(int n, int *ns, double limit) {
    std::vector<int> v(ns, ns + n);
    v = scale_slices_to_limit(v, limit);
    for (int i = 0; i < n; i++) ns[i] = v[i];
    return 0;
}

int orc_partition(const float *vals, int n, int nslices, float min_val, float max_val, float *centers, int32_t *to_bin) {
    Partition1D P = make_partition(std::vector<float>(vals, vals + n), nslices, min_val, max_val);
    std::copy(P.centers.begin(), P.centers.end(), centers);
    std::copy(P.to_bin.begin(), P.to_bin.end(), to_bin);
    return 0;
}

void *orc_build(const pt_params *p, const float *b, int h, int w, int c, const int32_t *allowed) {
    return build(*p, b, h, w, c, allowed);
}

void orc_free(void *t) { delete (Table *)t; }

// Learned-key table: the key of b ((h+7) x (w+7) x 8, computed by the GPU) indexes the table,
// with default slices over its 8 dimensions; exact RGB SSD for every distance; the GCK table
// built with the same parameters is the further table (its shuffle continues the rand() stream).
void *orc_build_keyed(const pt_params *p_in, const float *b, int h, int w, const float *key_hwc) {
    pt_params p = *p_in;
    if (p.grid_ndims < 0) p.grid_ndims = p.ndims;
    pt_params pk = p;
    pk.ndims = KEY_OUT;
    pk.nchroma = 0;
    pk.grid_ndims = KEY_OUT;
    pk.ntables = 1;
    pk.nslices_count = 0;
    g_key_hwc = key_hwc;
    Table *T = build(pk, b, h, w, 3, nullptr);
    g_key_hwc = nullptr;
    T->keyed = true;
    pt_params pg = p;
    pg.kcoherence = 0;
    pg.ntables = 1;
    T->extra.push_back(build(pg, b, h, w, 3, nullptr, T->rand_used));
    return T;
}

int orc_lookup_parallel_keyed(void *t, const float *a, int h, int w, const float *qkey, double *ann) {
    g_query_key = qkey;
    lookup_parallel(*(Table *)t, a, h, w, 3, ann, nullptr, nullptr);
    g_query_key = nullptr;
    return 0;
}

int orc_lookup_raster_keyed(void *t, const float *a, int h, int w, const float *qkey, double *ann) {
    g_query_key = qkey;
    lookup_raster(*(Table *)t, a, h, w, 3, ann, nullptr, nullptr);
    g_query_key = nullptr;
    return 0;
}

// The reference's table keeps a pointer to the caller's parameters (patchtable.h:2602), so
// parameters changed after construction apply to later lookups; this mirrors that.
int orc_set_params(void *t, const pt_params *p) {
    Table *T = (Table *)t;
    int g = T->p.grid_ndims;
    T->p = *p;
    T->p.grid_ndims = g;
    return 0;
}

int orc_shape(void *t, int *h, int *w, int *nd) {
    Table *T = (Table *)t;
    *h = T->h;
    *w = T->w;
    *nd = T->ndims;
    return 0;
}

int orc_info(void *t, int *G, int *nslices, int64_t *ncells, int *add_count, int *K) {
    Table *T = (Table *)t;
    *G = T->G;
    for (int i = 0; i < T->G; i++) nslices[i] = T->nslices[i];
    *ncells = T->ncells;
    *add_count = T->table_add_count;
    *K = T->K;
    return 0;
}

int orc_extra_info(void *t, int i, int *nslices, int64_t *ncells, int *add_count) {
    Table *T = (Table *)t;
    if (i < 0 || i >= (int)T->extra.size()) return -1;
    const Table *E = T->extra[i];
    for (int d = 0; d < E->G; d++) nslices[d] = E->nslices[d];
    *ncells = E->ncells;
    *add_count = E->table_add_count;
    return 0;
}

int orc_get_extra_table(void *t, int i, int32_t *out) {
    Table *T = (Table *)t;
    if (i < 0 || i >= (int)T->extra.size()) return -1;
    std::copy(T->extra[i]->table.begin(), T->extra[i]->table.end(), out);
    return 0;
}

int orc_get_descriptor(void *t, float *out) {
    Table *T = (Table *)t;
    std::copy(T->wh0.begin(), T->wh0.end(), out);
    return 0;
}

int orc_get_partition(void *t, int dim, float *mn, float *mx, float *centers, int32_t *to_bin) {
    Table *T = (Table *)t;
    const Partition1D &P = T->part[dim];
    *mn = P.min_val;
    *mx = P.max_val;
    std::copy(P.centers.begin(), P.centers.end(), centers);
    std::copy(P.to_bin.begin(), P.to_bin.end(), to_bin);
    return 0;
}

int orc_get_table(void *t, int32_t *out) {
    Table *T = (Table *)t;
    std::copy(T->table.begin(), T->table.end(), out);
    return 0;
}

int orc_get_populated(void *t, int32_t *out) {
    Table *T = (Table *)t;
    std::copy(T->table_pop.begin(), T->table_pop.end(), out);
    return 0;
}

int orc_get_kcoherence(void *t, int32_t *pos, float *dist) {
    Table *T = (Table *)t;
    std::copy(T->kpos.begin(), T->kpos.end(), pos);
    std::copy(T->kdist.begin(), T->kdist.end(), dist);
    return 0;
}

// PatchTableFixedN::save (patchtable.h:1574-1582) with the serialisation of patchtable.cpp:499-558:
// strings and vectors prefixed with their int32 length, Arrays with their sizes vector.
int orc_save(void *t, const char *path) {
    Table *T = (Table *)t;
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    auto raw = [&](const void *d, size_t n) { fwrite(d, 1, n, f); };
    auto i32 = [&](int32_t v) { raw(&v, 4); };
    auto str = [&](const char *s) { i32((int32_t)strlen(s)); raw(s, strlen(s)); };
    auto ivecw = [&](const std::vector<int> &v) { i32((int32_t)v.size()); raw(v.data(), v.size() * 4); };
    auto fvecw = [&](const fvec &v) { i32((int32_t)v.size()); raw(v.data(), v.size() * 4); };
    str("PatchTable");
    ivecw(T->nslices);
    raw(T->table.data(), T->table.size() * 4);
    ivecw({T->h, T->w, 3});
    raw(T->b0.data(), T->b0.size() * 4);
    ivecw({T->h + 7, T->w + 7, T->ndims});
    raw(T->wh0.data(), T->wh0.size() * 4);
    ivecw(T->nslices);
    str("PatchPartition");
    i32(T->G);
    for (const Partition1D &P : T->part) {
        str("PatchPartition1D");
        ivecw(P.to_bin);
        fvecw(P.centers);
        fvecw(P.lsizes);
        fvecw(P.rsizes);
        raw(&P.min_val, 4);
        raw(&P.max_val, 4);
        raw(&P.range, 4);
        i32(P.nslices);
        str("EndPatchPartition1D");
    }
    str("EndPatchPartition");
    str("EndPatchTable");
    return fclose(f) == 0 ? 0 : -1;
}

// Re-run a distance transform on a given populated table (tests of the DT variants).
// algo: 0 = prop BFS, 1 = prop separable, 2 = downsample, 3 = euclidean.
int orc_dt(int G, const int *nslices, int32_t *table, int algo, int dt_iters) {
    Table T;
    T.G = G;
    T.nslices.assign(nslices, nslices + G);
    T.stride.assign(G, 0);
    int64_t acc = 1;
    for (int i = G - 1; i >= 0; i--) {
        T.stride[i] = acc;
        acc *= nslices[i];
    }
    T.ncells = acc;
    T.table.assign(table, table + acc);
    if (algo == 0) dt_prop_bfs(T, dt_iters);
    else if (algo == 1) dt_prop_separable(T, dt_iters > 0 ? dt_iters + 1 : -1);
    else if (algo == 2) dt_downsample(T);
    else if (algo == 3) dt_euclidean(T);
    else return -1;
    std::copy(T.table.begin(), T.table.end(), table);
    return 0;
}

int orc_lookup_raster(void *t, const float *a, int h, int w, int c, double *ann, const double *prev, const float *sv) {
    lookup_raster(*(Table *)t, a, h, w, c, ann, prev, sv);
    return 0;
}

void orc_set_mfma_seen_skip(int on) { g_mfma_seen_skip = on; }

int orc_mfma_items(long *out, int n) {
    const int m = (int)g_mfma_items.size();
    for (int i = 0; i < n && i < m; i++) out[i] = g_mfma_items[i];
    return m;
}

int orc_ssd_counts(long *out) {
    for (int i = 0; i < 64; i++) out[i] = g_ssd_count[i];
    return 0;
}

int orc_lookup_parallel(void *t, const float *a, int h, int w, int c, double *ann, const double *prev, const float *sv) {
    lookup_parallel(*(Table *)t, a, h, w, c, ann, prev, sv);
    return 0;
}

float orc_desc_dist(const float *a, const float *b, int C) { return desc_dist(a, b, C); }

// Key planes (8 x (h+7) x (w+7), patch (x, y) at (y+7, x+7); other entries untouched).
int orc_key_forward(const float *W1, const float *B1, const float *W2, const float *B2, const float *img, int h, int w,
                    float *planes) {
    KeyNet net(W1, B1, W2, B2);
    const size_t plane = (size_t)(h + 7) * (w + 7);
    float xv[KEY_IN];
    double z[KEY_H], k[KEY_OUT];
    for (int y = 0; y + 8 <= h; y++)
        for (int x = 0; x + 8 <= w; x++) {
            key_patch(img, w, x, y, xv);
            net.forward(xv, z, k, true);
            for (int o = 0; o < KEY_OUT; o++) planes[o * plane + (size_t)(y + 7) * (w + 7) + x + 7] = (float)k[o];
        }
    return 0;
}

// Gradients of the mean triplet margin loss on squared key distances (anchor in a, positive and
// negative in b; pos[3t + role] = XY_TO_INT), in double.
int orc_key_gradients(const float *W1, const float *B1, const float *W2, const float *B2, const float *a,
                      const float *b, int w, const int32_t *pos, int nt, float margin, float *gW1, float *gb1,
                      float *gW2, float *gb2, float *loss) {
    KeyNet net(W1, B1, W2, B2);
    std::vector<double> G1((size_t)KEY_H * KEY_IN, 0.0), Gb1(KEY_H, 0.0), G2((size_t)KEY_OUT * KEY_H, 0.0), Gb2(KEY_OUT, 0.0);
    double L = 0.0;
    float xv[3][KEY_IN];
    double z[3][KEY_H], k[3][KEY_OUT];
    for (int t = 0; t < nt; t++) {
        for (int r = 0; r < 3; r++) {
            const int v = pos[3 * t + r];
            key_patch(r == 0 ? a : b, w, INT_TO_X(v), INT_TO_Y(v), xv[r]);
            net.forward(xv[r], z[r], k[r]);
        }
        double dap = 0, dan = 0;
        for (int o = 0; o < KEY_OUT; o++) {
            dap += (k[0][o] - k[1][o]) * (k[0][o] - k[1][o]);
            dan += (k[0][o] - k[2][o]) * (k[0][o] - k[2][o]);
        }
        const double l = dap - dan + margin;
        if (l <= 0) continue;
        L += l / nt;
        double dk[3][KEY_OUT];
        for (int o = 0; o < KEY_OUT; o++) {
            dk[0][o] = 2.0 * (k[2][o] - k[1][o]) / nt;
            dk[1][o] = -2.0 * (k[0][o] - k[1][o]) / nt;
            dk[2][o] = 2.0 * (k[0][o] - k[2][o]) / nt;
        }
        for (int r = 0; r < 3; r++) {
            for (int o = 0; o < KEY_OUT; o++) {
                Gb2[o] += dk[r][o];
                for (int j = 0; j < KEY_H; j++) G2[o * KEY_H + j] += dk[r][o] * (z[r][j] > 0 ? z[r][j] : 0.0);
            }
            for (int j = 0; j < KEY_H; j++) {
                if (z[r][j] <= 0) continue;
                double dh = 0;
                for (int o = 0; o < KEY_OUT; o++) dh += net.w2[o * KEY_H + j] * dk[r][o];
                Gb1[j] += dh;
                for (int i = 0; i < KEY_IN; i++) G1[j * KEY_IN + i] += dh * xv[r][i];
            }
        }
    }
    for (size_t i = 0; i < G1.size(); i++) gW1[i] = (float)G1[i];
    for (int i = 0; i < KEY_H; i++) gb1[i] = (float)Gb1[i];
    for (size_t i = 0; i < G2.size(); i++) gW2[i] = (float)G2[i];
    for (int i = 0; i < KEY_OUT; i++) gb2[i] = (float)Gb2[i];
    *loss = (float)L;
    return 0;
}

// Exact nearest neighbour in FP64 of each query patch (positions XY_TO_INT in patch coordinates)
// among every patch of b: sum over the 192 elements in row-major order with fused multiply-adds,
// ties to the smaller database index y*(bw-7)+x. pos_out is XY_TO_INT of the match.
int orc_brute_f64(const float *a, int aw, const int32_t *qpos, int nq, const float *b, int bh, int bw,
                  int32_t *pos_out, double *dist_out) {
    const int dew = bw - 7, deh = bh - 7;
    for (int q = 0; q < nq; q++) {
        const int qx = INT_TO_X(qpos[q]), qy = INT_TO_Y(qpos[q]);
        double qv[192];
        for (int k = 0; k < 192; k++) qv[k] = a[((size_t)(qy + k / 24) * aw + qx) * 3 + k % 24];
        double best = 1e300;
        int bidx = -1;
        for (int dy = 0; dy < deh; dy++)
            for (int dx = 0; dx < dew; dx++) {
                double acc = 0.0;
                for (int k = 0; k < 192; k++) {
                    const double d = qv[k] - (double)b[((size_t)(dy + k / 24) * bw + dx) * 3 + k % 24];
                    acc = std::fma(d, d, acc);
                }
                if (acc < best) {
                    best = acc;
                    bidx = dy * dew + dx;
                }
            }
        pos_out[q] = XY_TO_INT(bidx % dew, bidx / dew);
        dist_out[q] = best;
    }
    return 0;
}

int orc_key_round_weights(int on) {
    g_key_round_weights = on != 0;
    return 0;
}

// The mean triplet loss in double for double weights (the finite-difference side of the
// gradient check).
double orc_key_loss(const double *W1, const double *B1, const double *W2, const double *B2, const float *a,
                    const float *b, int w, const int32_t *pos, int nt, double margin) {
    KeyNet net(W1, B1, W2, B2);
    float xv[KEY_IN];
    double z[KEY_H], k[3][KEY_OUT], L = 0.0;
    for (int t = 0; t < nt; t++) {
        for (int r = 0; r < 3; r++) {
            const int v = pos[3 * t + r];
            key_patch(r == 0 ? a : b, w, INT_TO_X(v), INT_TO_Y(v), xv);
            net.forward(xv, z, k[r]);
        }
        double dap = 0, dan = 0;
        for (int o = 0; o < KEY_OUT; o++) {
            dap += (k[0][o] - k[1][o]) * (k[0][o] - k[1][o]);
            dan += (k[0][o] - k[2][o]) * (k[0][o] - k[2][o]);
        }
        const double l = dap - dan + margin;
        if (l > 0) L += l / nt;
    }
    return L;
}

// recompute_mean_dist<1> (patchtable_main.cpp:103-172, descriptor mode): float64 squared
// distance of the descriptors of every match (clamped), mean of the square roots.
double orc_recompute_mean_dist_desc(const float *a, int ah, int aw, const float *b, int bh, int bw, int C,
                                    const double *ann, double *dist_out) {
    double sum = 0.0;
    for (int ay = 0; ay < ah; ay++)
        for (int ax = 0; ax < aw; ax++) {
            const double *q = &ann[((size_t)ay * aw + ax) * 3];
            int bx = (int)q[0], by = (int)q[1];
            if (bx < 0) bx = 0;
            else if (bx >= bw) bx = bw - 1;
            if (by < 0) by = 0;
            else if (by >= bh) by = bh - 1;
            double dist = 0.0;
            for (int c = 0; c < C; c++) {
                double delta = (double)a[((size_t)ay * aw + ax) * C + c] - (double)b[((size_t)by * bw + bx) * C + c];
                dist += delta * delta;
            }
            if (dist_out) dist_out[(size_t)ay * aw + ax] = dist;
            sum += sqrt(dist);
        }
    return sum / double(aw * ah);
}

// Mean of sqrt(max(SSD,0)) over the field (get_mean_dist, patchtable_main.cpp:90-101).
double orc_mean_dist(const double *ann, long n) {
    double s = 0;
    for (long i = 0; i < n; i++) {
        double d = ann[i * 3 + 2];
        if (d < 0) d = 0;
        s += sqrt(d);
    }
    return s / double(n);
}

// recompute_mean_dist<0> (patchtable_main.cpp:103-172): float64 SSD of every match, clamped
// coordinates; optionally stores the recomputed SSD per patch in dist_out.
double orc_recompute_mean_dist(const float *a, int ah, int aw, const float *b, int bh, int bw, const double *ann, double *dist_out) {
    int ann_w = aw - 7, ann_h = ah - 7, bew = bw - 7, beh = bh - 7;
    double sum = 0.0;
    for (int ay = 0; ay < ann_h; ay++)
        for (int ax = 0; ax < ann_w; ax++) {
            const double *q = &ann[((size_t)ay * ann_w + ax) * 3];
            int bx = (int)q[0], by = (int)q[1];
            if (bx < 0) bx = 0;
            else if (bx >= bew) bx = bew - 1;
            if (by < 0) by = 0;
            else if (by >= beh) by = beh - 1;
            double dist = 0.0;
            for (int dy = 0; dy < 8; dy++)
                for (int dx = 0; dx < 8; dx++)
                    for (int c = 0; c < 3; c++) {
                        double delta = a[((size_t)(ay + dy) * aw + ax + dx) * 3 + c] - b[((size_t)(by + dy) * bw + bx + dx) * 3 + c];
                        dist += delta * delta;
                    }
            if (dist_out) dist_out[(size_t)ay * ann_w + ax] = dist;
            sum += sqrt(dist);
        }
    return sum / double(ann_w * ann_h);
}

}  // extern "C"
