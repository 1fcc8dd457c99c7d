// Distance-transform table fill on gfx950 (default dt_algo "prop", patchtable.h:2081-2165).
//
// The reference fills empty cells by a level-synchronous multi-source BFS whose edge labels
// only allow paths that move along dim 0 first, then dim 1, ... (PropEdge/PropElement,
// patchtable.h:1447-1468; next_label built at :2100-2120), where the first claim in list order
// wins. The initial list is the filled cells in increasing linear index, and the list order of
// every later level is lexicographic in (source index, edge sequence) of the claiming path.
// Induction over the levels shows every empty cell c is claimed at level d(c) = its L1 distance
// (in cell units) to the nearest filled cell, by the nearest filled cell of smallest linear
// index: every prefix of that source's dimension-ordered path to c is claimed by the same
// source through that path, and any other claimant at that level comes from a source of larger
// index. The fill is therefore a separable L1 distance transform over the cell lattice that
// minimises the pair (distance, source index) -- computed here as one forward and one backward
// sweep along each dimension. dt_iters > 0 (the reference stops after dt_iters+1 levels) keeps
// cells farther than dt_iters+1 empty, to be filled by fill_zeros (patchtable.h:2063-2079).
//
// Keys are 64-bit: distance in the high word, source cell index in the low word; empty = ~0.
#include "pt_device.h"
#include "pt_kernels.h"

#include <algorithm>
#include <vector>

namespace {

constexpr uint64_t KEY_INF = ~0ull;
constexpr uint64_t KEY_STEP = 1ull << 32;

__global__ void k_dt_init(const int32_t *__restrict__ table, int64_t n, uint64_t *__restrict__ key) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    key[c] = table[c] != PT_TABLE_UNUSED ? (uint64_t)c : KEY_INF;
}

// One line per lane along a dimension of size len and stride st (in cells).
__global__ void k_dt_line(uint64_t *__restrict__ key, int64_t nlines, int len, int64_t st) {
    int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    int64_t o = l / st, i = l - o * st;
    uint64_t *p = key + o * (int64_t)len * st + i;
    uint64_t run = KEY_INF;
    for (int t = 0; t < len; t++) {
        uint64_t k = p[(int64_t)t * st];
        if (run != KEY_INF) run += KEY_STEP;
        if (k < run) run = k;
        p[(int64_t)t * st] = run;
    }
    run = KEY_INF;
    for (int t = len - 1; t >= 0; t--) {
        uint64_t k = p[(int64_t)t * st];
        if (run != KEY_INF) run += KEY_STEP;
        if (k < run) run = k;
        p[(int64_t)t * st] = run;
    }
}

// The same two sweeps with the line held in registers (len <= LMAX): one read and one write
// per cell instead of two of each.
// With FINAL (the last dimension's pass) the lane writes the table values of k_dt_final for its
// line instead of the keys.
template <int LMAX, bool FINAL>
__global__ __launch_bounds__(64) void k_dt_line_reg(uint64_t *__restrict__ key, int64_t nlines, int len, int64_t st,
                                                    int32_t *__restrict__ table, int64_t max_dist, int32_t fill) {
    int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    int64_t o = l / st, i = l - o * st;
    uint64_t *p = key + o * (int64_t)len * st + i;
    uint64_t v[LMAX];
#pragma unroll
    for (int t = 0; t < LMAX; t++)
        if (t < len) v[t] = p[(int64_t)t * st];
    uint64_t run = KEY_INF;
#pragma unroll
    for (int t = 0; t < LMAX; t++)
        if (t < len) {
            if (run != KEY_INF) run += KEY_STEP;
            if (v[t] < run) run = v[t];
            v[t] = run;
        }
    run = KEY_INF;
#pragma unroll
    for (int t = LMAX - 1; t >= 0; t--)
        if (t < len) {
            if (run != KEY_INF) run += KEY_STEP;
            if (v[t] < run) run = v[t];
            v[t] = run;
        }
    if (FINAL) {
        // as k_dt_final: source cells keep their own value, so reading table[src] is safe while
        // other lanes rewrite other cells
        const int64_t c0 = o * (int64_t)len * st + i;
#pragma unroll
        for (int t = 0; t < LMAX; t++)
            if (t < len) {
                const uint64_t k = v[t];
                const int64_t c = c0 + (int64_t)t * st;
                int32_t val;
                if (k == KEY_INF || (max_dist >= 0 && (int64_t)(k >> 32) > max_dist)) val = fill;
                else val = table[(int64_t)(k & 0xffffffffull)];
                if (val != table[c]) table[c] = val;
            }
        return;
    }
#pragma unroll
    for (int t = 0; t < LMAX; t++)
        if (t < len) p[(int64_t)t * st] = v[t];
}

// The innermost m dimensions at once. For each index of the outer dimensions their cells form a
// contiguous slab; a block stages `spb` consecutive slabs (<= PT_DT_SLAB cells) in LDS with
// coalesced loads, runs both sweeps along each slab dimension in turn, innermost first (the
// order of the separate passes), and writes the keys back once. With `table` (the first pass)
// the keys are made from the table as k_dt_init makes them.
struct SlabDims {
    int m;
    int len[PT_MAX_DIMS];   // slab dimensions, innermost first
    int st[PT_MAX_DIMS];    // their strides in cells
};
constexpr int PT_DT_SLAB = 6144;   // cells of the staged slabs: 48 KB of keys

__global__ __launch_bounds__(256) void k_dt_slab(uint64_t *__restrict__ key, int64_t ncells, int P, int spb, SlabDims S,
                                                 const int32_t *__restrict__ table) {
    extern __shared__ uint64_t sh[];
    const int64_t c0 = (int64_t)blockIdx.x * spb * P;
    const int total = (int)(ncells - c0 < (int64_t)spb * P ? ncells - c0 : (int64_t)spb * P);
    uint64_t *g = key + c0;
    if (table) {
        for (int i = threadIdx.x; i < total; i += blockDim.x)
            sh[i] = table[c0 + i] != PT_TABLE_UNUSED ? (uint64_t)(c0 + i) : KEY_INF;
    } else {
        for (int i = threadIdx.x; i < total; i += blockDim.x) sh[i] = g[i];
    }
    __syncthreads();
    for (int k = 0; k < S.m; k++) {
        const int len = S.len[k], st = S.st[k];
        const int nl = total / len;   // lines of dimension k in the staged slabs
        for (int l = threadIdx.x; l < nl; l += blockDim.x) {
            uint64_t *p = sh + (l / st) * (len * st) + (l % st);
            uint64_t run = KEY_INF;
            for (int t = 0; t < len; t++) {
                if (run != KEY_INF) run += KEY_STEP;
                const uint64_t v = p[t * st];
                if (v < run) run = v;
                p[t * st] = run;
            }
            run = KEY_INF;
            for (int t = len - 1; t >= 0; t--) {
                if (run != KEY_INF) run += KEY_STEP;
                const uint64_t v = p[t * st];
                if (v < run) run = v;
                p[t * st] = run;
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < total; i += blockDim.x) g[i] = sh[i];
}

// table[c] = value of the winning source, or `fill` for cells beyond max_dist / unreached.
// Source cells keep their own value, so reading table[src] while writing other cells is safe.
__global__ void k_dt_final(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ table, int64_t max_dist,
                           int32_t fill) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    uint64_t k = key[c];
    int32_t v;
    if (k == KEY_INF || (max_dist >= 0 && (int64_t)(k >> 32) > max_dist)) v = fill;
    else v = table[(int64_t)(k & 0xffffffffull)];
    if (v != table[c]) table[c] = v;
}

// fill_zeros (patchtable.h:2063-2079): every empty cell takes `fill`.
__global__ void k_fill_unused(int32_t *__restrict__ table, int64_t n, int32_t fill) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n && table[c] == PT_TABLE_UNUSED) table[c] = fill;
}

// ---- DOWNSAMPLE (do_dt_downsample, patchtable.h:2339-2361; dt_downsample.h) ----------------
struct LevelShape {
    int G;
    int hs[PT_MAX_DIMS];      // fine sizes
    int ls[PT_MAX_DIMS];      // coarse sizes ((hs+1)/2)
    int64_t hst[PT_MAX_DIMS]; // fine strides
    int64_t lst[PT_MAX_DIMS]; // coarse strides
};

// A coarse cell takes the first filled child (no TABLE_HI_MASK bit) in nested-loop order with
// dim 0 outermost, else TABLE_UNUSED.
__global__ void k_downsample(const int32_t *__restrict__ hi, int32_t *__restrict__ lo, int64_t nlo, LevelShape S) {
    int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nlo) return;
    int64_t base = 0;
    int cnt[PT_MAX_DIMS];
    int64_t nchild = 1;
    for (int i = 0; i < S.G; i++) {
        int lc = (int)((li / S.lst[i]) % S.ls[i]);
        int c0 = 2 * lc;
        int c1 = 2 * (lc + 1) < S.hs[i] ? 2 * (lc + 1) : S.hs[i];
        cnt[i] = c1 - c0;
        base += (int64_t)c0 * S.hst[i];
        nchild *= cnt[i];
    }
    int32_t out = PT_TABLE_UNUSED;
    for (int64_t c = 0; c < nchild; c++) {
        int64_t rem = c, off = 0;
        for (int i = S.G - 1; i >= 0; i--) {
            int ci = (int)(rem % cnt[i]);
            rem /= cnt[i];
            off += (int64_t)ci * S.hst[i];
        }
        int32_t v = hi[base + off];
        if (!(v & (1 << 30)) && v != PT_TABLE_UNUSED) {
            out = v;
            break;
        }
    }
    lo[li] = out;
}

// An empty fine cell copies its coarse parent.
__global__ void k_upsample(int32_t *__restrict__ hi, const int32_t *__restrict__ lo, int64_t nhi, LevelShape S) {
    int64_t hidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (hidx >= nhi) return;
    if (hi[hidx] != PT_TABLE_UNUSED) return;
    int64_t lidx = 0;
    for (int i = 0; i < S.G; i++) lidx += (int64_t)(((hidx / S.hst[i]) % S.hs[i]) / 2) * S.lst[i];
    hi[hidx] = lo[lidx];
}

// ---- EUCLIDEAN (euclidean_dt + do_dt_euclidean, patchtable.h:4244-4352) ---------------------
// Felzenszwalb-Huttenlocher lower envelope along one dimension per pass, dims in order
// 0..G-1, in the reference's float arithmetic; the carried table value is updated in place
// exactly as the reference's index_array (idx[q] = idx[v[k]]). Per-line work arrays live in
// global memory at the positions of the line's cells.
__global__ void k_eucl_line(int32_t *__restrict__ idx, float *__restrict__ f, float *__restrict__ dist,
                            int32_t *__restrict__ vb, float *__restrict__ zb, float *__restrict__ zlast, int64_t nlines,
                            int n, int64_t st) {
    int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    int64_t o = l / st, i = l - o * st;
    const int64_t base = o * (int64_t)n * st + i;
    const float INF = 1E20;
#define F_(q) f[base + (int64_t)(q) * st]
#define V_(k) vb[base + (int64_t)(k) * st]
#define Z_(k) (*((k) < n ? &zb[base + (int64_t)(k) * st] : &zlast[l]))
    int k = 0;
    V_(0) = 0;
    Z_(0) = -INF;
    Z_(1) = +INF;
    for (int q = 1; q <= n - 1; q++) {
        int vk = V_(k);
        float s = ((F_(q) + q * q) - (F_(vk) + vk * vk)) / (2 * q - 2 * vk);
        while (s <= Z_(k)) {
            k--;
            vk = V_(k);
            s = ((F_(q) + q * q) - (F_(vk) + vk * vk)) / (2 * q - 2 * vk);
        }
        k++;
        V_(k) = q;
        Z_(k) = s;
        Z_(k + 1) = +INF;
    }
    k = 0;
    for (int q = 0; q <= n - 1; q++) {
        while (Z_(k + 1) < q) k++;
        int vk = V_(k);
        dist[base + (int64_t)q * st] = (q - vk) * (q - vk) + F_(vk);
        idx[base + (int64_t)q * st] = idx[base + (int64_t)vk * st];
    }
    for (int q = 0; q <= n - 1; q++) F_(q) = dist[base + (int64_t)q * st];
#undef F_
#undef V_
#undef Z_
}

__global__ void k_eucl_init(const int32_t *__restrict__ table, int64_t n, float *__restrict__ f) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) f[c] = table[c] == PT_TABLE_UNUSED ? (float)1E20 : 0.0f;
}

}  // namespace

int pt_fill_unused_launch(int32_t *table, int64_t ncells, int32_t fill, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_unused, dim3(pt_blocks(ncells, 256)), dim3(256), 0, s, table, ncells, fill);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_dt_downsample_launch(int32_t *table, int G, const int *nslices, PtScratch &sc, hipStream_t s) {
    std::vector<std::vector<int>> sizes(1, std::vector<int>(nslices, nslices + G));
    auto all_one = [](const std::vector<int> &v) {
        for (int x : v)
            if (x != 1) return false;
        return true;
    };
    while (!all_one(sizes.back())) {
        std::vector<int> ls(G);
        for (int i = 0; i < G; i++) ls[i] = (sizes.back()[i] + 1) / 2;
        sizes.push_back(ls);
    }
    const int L = (int)sizes.size();
    std::vector<int32_t *> lvl(L, nullptr);
    std::vector<int64_t> cnt(L);
    lvl[0] = table;
    for (int k = 0; k < L; k++) {
        int64_t c = 1;
        for (int v : sizes[k]) c *= v;
        cnt[k] = c;
        if (k > 0) {
            lvl[k] = (int32_t *)sc.get(50 + k, (size_t)c * 4);
            if (!lvl[k]) return pt_fail(PT_E_NOMEM, "downsample pyramid allocation failed");
        }
    }
    auto shape = [&](int k) {   // fine level k-1 -> coarse level k
        LevelShape S;
        S.G = G;
        int64_t ah = 1, al = 1;
        for (int i = G - 1; i >= 0; i--) {
            S.hs[i] = sizes[k - 1][i];
            S.ls[i] = sizes[k][i];
            S.hst[i] = ah;
            S.lst[i] = al;
            ah *= S.hs[i];
            al *= S.ls[i];
        }
        return S;
    };
    for (int k = 1; k < L; k++)
        hipLaunchKernelGGL(k_downsample, dim3(pt_blocks(cnt[k], 256)), dim3(256), 0, s, lvl[k - 1], lvl[k], cnt[k], shape(k));
    for (int k = L - 1; k >= 1; k--)
        hipLaunchKernelGGL(k_upsample, dim3(pt_blocks(cnt[k - 1], 256)), dim3(256), 0, s, lvl[k - 1], lvl[k], cnt[k - 1],
                           shape(k));
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_dt_euclidean_launch(int32_t *table, int G, const int *nslices, int64_t ncells, float *f, hipStream_t s) {
    // f: ncells floats (the reference's table_dist); work arrays below.
    float *dist = nullptr, *zb = nullptr, *zlast = nullptr;
    int32_t *vb = nullptr;
    int64_t maxlines = 0;
    for (int d = 0; d < G; d++) maxlines = std::max<int64_t>(maxlines, ncells / nslices[d]);
    if (hipMalloc(&dist, ncells * 4) != hipSuccess || hipMalloc(&zb, ncells * 4) != hipSuccess ||
        hipMalloc(&vb, ncells * 4) != hipSuccess || hipMalloc(&zlast, maxlines * 4) != hipSuccess) {
        (void)hipFree(dist);
        (void)hipFree(zb);
        (void)hipFree(vb);
        (void)hipFree(zlast);
        return pt_fail(PT_E_NOMEM, "euclidean dt work arrays allocation failed");
    }
    hipLaunchKernelGGL(k_eucl_init, dim3(pt_blocks(ncells, 256)), dim3(256), 0, s, table, ncells, f);
    std::vector<int64_t> stride(G);
    int64_t acc = 1;
    for (int d = G - 1; d >= 0; d--) {
        stride[d] = acc;
        acc *= nslices[d];
    }
    for (int d = 0; d < G; d++) {
        int64_t nlines = ncells / nslices[d];
        hipLaunchKernelGGL(k_eucl_line, dim3(pt_blocks(nlines, 64)), dim3(64), 0, s, table, f, dist, vb, zb, zlast,
                           nlines, nslices[d], stride[d]);
    }
    hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(dist);
    (void)hipFree(zb);
    (void)hipFree(vb);
    (void)hipFree(zlast);
    if (e != hipSuccess) return pt_failf(PT_E_HIP, "euclidean dt failed: %s", hipGetErrorString(e));
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_dt_prop_launch(int32_t *table, int G, const int *nslices, int64_t ncells, int dt_iters, int32_t fill,
                      uint64_t *key, hipStream_t s) {
    const int TB = 256;
    const int64_t max_dist = dt_iters > 0 ? (int64_t)dt_iters + 1 : -1;
    // passes run d = G-1 .. 0. The innermost dimensions whose cells fit PT_DT_SLAB go in one
    // slab pass, which also makes the keys from the table; the last (strided, in registers)
    // pass writes the table when dimension 0 is outside the slab and short.
    SlabDims S{};
    int64_t P = 1;
    while (S.m < G && P * nslices[G - 1 - S.m] <= PT_DT_SLAB) {
        S.len[S.m] = nslices[G - 1 - S.m];
        S.st[S.m] = (int)P;
        P *= nslices[G - 1 - S.m];
        S.m++;
    }
    int64_t st0 = 1;   // stride of dimension 0
    for (int d = 1; d < G; d++) st0 *= nslices[d];
    const bool final_fused = st0 > 1 && nslices[0] <= 64 && S.m < G;
    int64_t st = 1;
    int dstart = G - 1;
    if (S.m > 0) {
        const int spb = (int)(PT_DT_SLAB / P);
        const int64_t nblk = (ncells + (int64_t)spb * P - 1) / ((int64_t)spb * P);
        hipLaunchKernelGGL(k_dt_slab, dim3((unsigned)nblk), dim3(256), (size_t)spb * P * 8, s, key, ncells, (int)P, spb,
                           S, (const int32_t *)table);
        st = P;
        dstart = G - 1 - S.m;
    } else {
        hipLaunchKernelGGL(k_dt_init, dim3(pt_blocks(ncells, TB)), dim3(TB), 0, s, table, ncells, key);
    }
    for (int d = dstart; d >= 0; d--) {
        const int64_t nlines = ncells / nslices[d];
        const int len = nslices[d];
        const dim3 g64(pt_blocks(nlines, 64));
        if (d == 0 && final_fused) {
            if (len <= 32)
                hipLaunchKernelGGL((k_dt_line_reg<32, true>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                                   max_dist, fill);
            else
                hipLaunchKernelGGL((k_dt_line_reg<64, true>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                                   max_dist, fill);
        } else if (len <= 32) {
            hipLaunchKernelGGL((k_dt_line_reg<32, false>), g64, dim3(64), 0, s, key, nlines, len, st, table, max_dist,
                               fill);
        } else if (len <= 64) {
            hipLaunchKernelGGL((k_dt_line_reg<64, false>), g64, dim3(64), 0, s, key, nlines, len, st, table, max_dist,
                               fill);
        } else {
            hipLaunchKernelGGL(k_dt_line, g64, dim3(64), 0, s, key, nlines, len, st);
        }
        st *= len;
    }
    if (!final_fused)
        hipLaunchKernelGGL(k_dt_final, dim3(pt_blocks(ncells, TB)), dim3(TB), 0, s, key, ncells, table, max_dist, fill);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
