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
// Keys order the pair (distance, source index) as one unsigned integer, empty = all ones:
//   32-bit keys (the default): distance << 27 | source, for tables of at most 2^27 cells. A key
//     holds distances up to K32_DMAX = 30; a step that would go beyond drops the key (it becomes
//     empty). Distances only grow along a path, so the result is the exact fill for every cell
//     whose distance is at most 30. A cell left empty at the end (distance above 30, or a table
//     with no filled cell) is not written: the last pass raises a flag, and the 64-bit passes then
//     first restore the cells the last pass wrote (a cell is a source exactly when its key's
//     distance is 0) and compute the whole fill again. The build reads the flag at its final
//     synchronisation and runs them only when it is set; without a host copy of the flag they are
//     launched behind it on the device and return at once when it is clear.
//   64-bit keys: distance << 32 | source, with no distance limit (the fallback, and the path for
//     lattices whose shape the 32-bit passes do not cover).
// Keys of both widths live in one buffer of 8-byte words; the flag is the int after it.
#include "pt_device.h"
#include "pt_kernels.h"

#include <algorithm>
#include <vector>

namespace {

template <typename K> struct DtKey;
template <> struct DtKey<uint64_t> {
    static constexpr uint64_t INF = ~0ull;
    static __device__ __forceinline__ uint64_t step(uint64_t r) { return r == INF ? r : r + (1ull << 32); }
    static __device__ __forceinline__ int64_t src(uint64_t k) { return (int64_t)(k & 0xffffffffull); }
    static __device__ __forceinline__ int64_t dist(uint64_t k) { return (int64_t)(k >> 32); }
};
constexpr int K32_SHIFT = 27;
constexpr uint32_t K32_DMAX = 30;
template <> struct DtKey<uint32_t> {
    static constexpr uint32_t INF = ~0u;
    // INF >> 27 is 31, so INF stays INF
    static __device__ __forceinline__ uint32_t step(uint32_t r) {
        return (r >> K32_SHIFT) >= K32_DMAX ? INF : r + (1u << K32_SHIFT);
    }
    static __device__ __forceinline__ int64_t src(uint32_t k) { return (int64_t)(k & ((1u << K32_SHIFT) - 1)); }
    static __device__ __forceinline__ int64_t dist(uint32_t k) { return (int64_t)(k >> K32_SHIFT); }
};

// run_if: null, or a device flag; the kernel does nothing while the flag is 0.
__device__ __forceinline__ bool dt_skip(const int *run_if) { return run_if && *(volatile const int *)run_if == 0; }

__global__ void k_dt_init(const int32_t *__restrict__ table, int64_t n, uint64_t *__restrict__ key, const int *run_if) {
    if (dt_skip(run_if)) return;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        key[c] = table[c] != PT_TABLE_UNUSED ? (uint64_t)c : DtKey<uint64_t>::INF;
}

// One line per lane along a dimension of size len and stride st (in cells), lines of any length.
__global__ void k_dt_line(uint64_t *__restrict__ key, int64_t nlines, int len, int64_t st, const int *run_if) {
    if (dt_skip(run_if)) return;
    typedef DtKey<uint64_t> D;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < nlines; l += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = l / st, i = l - o * st;
        uint64_t *p = key + o * (int64_t)len * st + i;
        uint64_t run = D::INF;
        for (int t = 0; t < len; t++) {
            uint64_t k = p[(int64_t)t * st];
            run = D::step(run);
            if (k < run) run = k;
            p[(int64_t)t * st] = run;
        }
        run = D::INF;
        for (int t = len - 1; t >= 0; t--) {
            uint64_t k = p[(int64_t)t * st];
            run = D::step(run);
            if (k < run) run = k;
            p[(int64_t)t * st] = run;
        }
    }
}

// Table value of cell c from its final key; returns false when the cell is not written: a
// source (it keeps its own value) or, with 32-bit keys, a cell the keys left empty (the flag is
// raised unless a distance cap at or below K32_DMAX makes it a `fill` cell anyway).
template <typename K>
__device__ __forceinline__ bool dt_value(K k, int64_t c, const int32_t *table, int64_t max_dist, int32_t fill,
                                         int *flag, int32_t &val) {
    typedef DtKey<K> D;
    if (k == D::INF) {
        if constexpr (sizeof(K) == 4) {
            if (!(max_dist >= 0 && max_dist <= (int64_t)K32_DMAX)) {
                *flag = 1;
                return false;
            }
        }
        val = fill;
        return true;
    }
    if (max_dist >= 0 && D::dist(k) > max_dist) {
        val = fill;
        return true;
    }
    const int64_t src = D::src(k);
    if (src == c) return false;
    val = table[src];
    return true;
}

// Both sweeps over a line of len <= LMAX keys held in registers.
template <typename K, int LMAX>
__device__ __forceinline__ void sweep_regs(K (&v)[LMAX], int len) {
    typedef DtKey<K> D;
    K run = D::INF;
#pragma unroll
    for (int t = 0; t < LMAX; t++)
        if (t < len) {
            run = D::step(run);
            if (v[t] < run) run = v[t];
            v[t] = run;
        }
    run = D::INF;
#pragma unroll
    for (int t = LMAX - 1; t >= 0; t--)
        if (t < len) {
            run = D::step(run);
            if (v[t] < run) run = v[t];
            v[t] = run;
        }
}

// The two sweeps with the line held in registers (len <= LMAX): one read and one write per cell.
// With FINAL (the last dimension's pass) the lane writes its line's table values instead of the
// keys: source cells are never written, so reading table[src] is safe while other lanes write.
template <typename K, int LMAX, bool FINAL>
__global__ __launch_bounds__(64) void k_dt_line_reg(K *__restrict__ key, int64_t nlines, int len, int64_t st,
                                                    int32_t *__restrict__ table, int64_t max_dist, int32_t fill,
                                                    int *flag, const int *run_if) {
    if (dt_skip(run_if)) return;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < nlines; l += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = l / st, i = l - o * st;
        const int64_t c0 = o * (int64_t)len * st + i;
        K *p = key + c0;
        K v[LMAX];
#pragma unroll
        for (int t = 0; t < LMAX; t++)
            if (t < len) v[t] = p[(int64_t)t * st];
        sweep_regs<K, LMAX>(v, len);
        if (FINAL) {
#pragma unroll
            for (int t = 0; t < LMAX; t++)
                if (t < len) {
                    const int64_t c = c0 + (int64_t)t * st;
                    int32_t val;
                    if (dt_value<K>(v[t], c, table, max_dist, fill, flag, val)) table[c] = val;
                }
        } else {
#pragma unroll
            for (int t = 0; t < LMAX; t++)
                if (t < len) p[(int64_t)t * st] = v[t];
        }
    }
}

// The innermost m dimensions at once. For each index of the outer dimensions their cells form a
// contiguous slab; a block stages `spb` consecutive slabs (P cells each) in LDS with coalesced
// loads, runs both sweeps along each slab dimension in turn, innermost first (the order of the
// separate passes), and writes the keys back once. The keys are made from the table (cell c:
// c if filled, else empty), as k_dt_init makes them.
struct SlabDims {
    int m;
    int len[PT_MAX_DIMS];   // slab dimensions, innermost first
    int st[PT_MAX_DIMS];    // their strides in cells
};
constexpr int PT_DT_SLAB = 6144;   // cells of the staged slabs (24 or 48 KB of keys)

template <typename K>
__global__ __launch_bounds__(256) void k_dt_slab(K *__restrict__ key, int64_t ncells, int P, int spb, SlabDims S,
                                                 const int32_t *__restrict__ table, const int *run_if) {
    if (dt_skip(run_if)) return;
    typedef DtKey<K> D;
    extern __shared__ __align__(8) unsigned char dt_lds[];
    K *sh = (K *)dt_lds;
    const int64_t per = (int64_t)spb * P;
    const int64_t nblk = (ncells + per - 1) / per;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int64_t c0 = blk * per;
        const int total = (int)(ncells - c0 < per ? ncells - c0 : per);
        for (int i = threadIdx.x; i < total; i += blockDim.x)
            sh[i] = table[c0 + i] != PT_TABLE_UNUSED ? (K)(c0 + i) : D::INF;
        __syncthreads();
        for (int k = 0; k < S.m; k++) {
            const int len = S.len[k], st = S.st[k];
            const int nl = total / len;   // lines of dimension k in the staged slabs
            for (int l = threadIdx.x; l < nl; l += blockDim.x) {
                K *p = sh + (l / st) * (len * st) + (l % st);
                K run = D::INF;
                for (int t = 0; t < len; t++) {
                    run = D::step(run);
                    const K v = p[t * st];
                    if (v < run) run = v;
                    p[t * st] = run;
                }
                run = D::INF;
                for (int t = len - 1; t >= 0; t--) {
                    run = D::step(run);
                    const K v = p[t * st];
                    if (v < run) run = v;
                    p[t * st] = run;
                }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < total; i += blockDim.x) key[c0 + i] = sh[i];
        __syncthreads();
    }
}

// Writes the table from 64-bit keys (when the last pass is not a register pass).
__global__ void k_dt_final(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ table, int64_t max_dist,
                           int32_t fill, const int *run_if) {
    if (dt_skip(run_if)) return;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int32_t val;
        if (dt_value<uint64_t>(key[c], c, table, max_dist, fill, nullptr, val)) table[c] = val;
    }
}

// Before the 64-bit fallback: every cell the 32-bit last pass may have written (the cells that are
// not sources: their key's distance is not 0) is empty again.
__global__ void k_dt_restore(const uint32_t *__restrict__ key, int64_t n, int32_t *__restrict__ table,
                             const int *run_if) {
    if (dt_skip(run_if)) return;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        if ((key[c] >> K32_SHIFT) != 0) table[c] = PT_TABLE_UNUSED;
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

namespace {

// The innermost dimensions whose cells fit `cap` cells, innermost first.
SlabDims slab_dims(int G, const int *nslices, int64_t cap, int64_t &P) {
    SlabDims S{};
    P = 1;
    while (S.m < G && P * nslices[G - 1 - S.m] <= cap) {
        S.len[S.m] = nslices[G - 1 - S.m];
        S.st[S.m] = (int)P;
        P *= nslices[G - 1 - S.m];
        S.m++;
    }
    return S;
}

// One sweep pass per dimension outside the slab, the last one (dimension 0) writing the table.
// grid_cap > 0 bounds the grids (the fallback's launches, which usually return at once).
template <typename K>
void dt_passes(int32_t *table, int G, const int *nslices, int64_t ncells, int64_t max_dist, int32_t fill, K *key,
               const SlabDims &S, int64_t P, int *flag, const int *run_if, unsigned grid_cap, hipStream_t s) {
    auto cap = [&](unsigned g) { return grid_cap && g > grid_cap ? grid_cap : g; };
    const int64_t st0 = ncells / nslices[0];
    const bool final_fused = st0 > 1 && nslices[0] <= 64 && S.m < G;
    int64_t st = 1;
    int dstart = G - 1;
    if (S.m > 0) {
        const int spb = (int)(PT_DT_SLAB / P);
        const int64_t nblk = (ncells + (int64_t)spb * P - 1) / ((int64_t)spb * P);
        hipLaunchKernelGGL(k_dt_slab<K>, dim3(cap((unsigned)nblk)), dim3(256), (size_t)spb * P * sizeof(K), s, key,
                           ncells, (int)P, spb, S, (const int32_t *)table, run_if);
        st = P;
        dstart = G - 1 - S.m;
    } else {
        // only 64-bit keys get here
        hipLaunchKernelGGL(k_dt_init, dim3(cap(pt_blocks(ncells, 256))), dim3(256), 0, s, table, ncells,
                           (uint64_t *)key, run_if);
    }
    for (int d = dstart; d >= 0; d--) {
        const int64_t nlines = ncells / nslices[d];
        const int len = nslices[d];
        const dim3 g64(cap(pt_blocks(nlines, 64)));
        if (d == 0 && final_fused) {
            if (len <= 32)
                hipLaunchKernelGGL((k_dt_line_reg<K, 32, true>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                                   max_dist, fill, flag, run_if);
            else
                hipLaunchKernelGGL((k_dt_line_reg<K, 64, true>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                                   max_dist, fill, flag, run_if);
        } else if (len <= 32) {
            hipLaunchKernelGGL((k_dt_line_reg<K, 32, false>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                               max_dist, fill, flag, run_if);
        } else if (len <= 64) {
            hipLaunchKernelGGL((k_dt_line_reg<K, 64, false>), g64, dim3(64), 0, s, key, nlines, len, st, table,
                               max_dist, fill, flag, run_if);
        } else {
            // only 64-bit keys get here
            hipLaunchKernelGGL(k_dt_line, g64, dim3(64), 0, s, (uint64_t *)key, nlines, len, st, run_if);
        }
        st *= len;
    }
    if (!final_fused)   // only 64-bit keys get here
        hipLaunchKernelGGL(k_dt_final, dim3(cap(pt_blocks(ncells, 256))), dim3(256), 0, s, (const uint64_t *)key,
                           ncells, table, max_dist, fill, run_if);
}

// The 32-bit passes cover: at most 2^27 cells, a slab of at least one and fewer than all
// dimensions, and the other dimensions at most 64 long (register passes).
bool dt_narrow(int G, const int *nslices, int64_t ncells) {
    if (ncells > ((int64_t)1 << K32_SHIFT) || G < 2) return false;
    int64_t P;
    const SlabDims S = slab_dims(G, nslices, PT_DT_SLAB, P);
    if (S.m == 0 || S.m == G) return false;
    for (int d = 0; d < G - S.m; d++)
        if (nslices[d] > 64) return false;
    return true;
}

}  // namespace

int pt_dt_prop_launch(int32_t *table, int G, const int *nslices, int64_t ncells, int dt_iters, int32_t fill,
                      uint64_t *key, hipStream_t s, int *flag_host) {
    const int64_t max_dist = dt_iters > 0 ? (int64_t)dt_iters + 1 : -1;
    int *flag = (int *)(key + ncells);
    int64_t P;
    const SlabDims S = slab_dims(G, nslices, PT_DT_SLAB, P);
    if (!dt_narrow(G, nslices, ncells)) {
        dt_passes<uint64_t>(table, G, nslices, ncells, max_dist, fill, key, S, P, nullptr, nullptr, 0, s);
        if (flag_host) *flag_host = 0;
        PT_LAUNCH_CHECK();
        return PT_OK;
    }
    if (int rc = pt_fill_launch(flag, 1, 0, s)) return rc;
    dt_passes<uint32_t>(table, G, nslices, ncells, max_dist, fill, (uint32_t *)key, S, P, flag, nullptr, 0, s);
    if (flag_host) {
        // the caller reads the flag once the stream has synchronised and, when it is set, runs
        // pt_dt_prop_fallback_launch
        PT_HIP_TRY(hipMemcpyAsync(flag_host, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    } else {
        // the fallback on the device: its launches return at once unless the flag is set
        const unsigned FB_GRID = 512;
        hipLaunchKernelGGL(k_dt_restore, dim3(FB_GRID), dim3(256), 0, s, (const uint32_t *)key, ncells, table,
                           (const int *)flag);
        dt_passes<uint64_t>(table, G, nslices, ncells, max_dist, fill, key, S, P, nullptr, flag, FB_GRID, s);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_dt_prop_fallback_launch(int32_t *table, int G, const int *nslices, int64_t ncells, int dt_iters, int32_t fill,
                               uint64_t *key, hipStream_t s) {
    const int64_t max_dist = dt_iters > 0 ? (int64_t)dt_iters + 1 : -1;
    int64_t P;
    const SlabDims S = slab_dims(G, nslices, PT_DT_SLAB, P);
    hipLaunchKernelGGL(k_dt_restore, dim3(pt_blocks(ncells, 256)), dim3(256), 0, s, (const uint32_t *)key, ncells, table,
                       nullptr);
    dt_passes<uint64_t>(table, G, nslices, ncells, max_dist, fill, key, S, P, nullptr, nullptr, 0, s);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
