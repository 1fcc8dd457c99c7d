// Quantiser reductions on gfx950: per-dimension min/max over the database patches
// (set_min_max, patchtable.h:1758-1778) and the order statistics that place the slice centres
// (PatchPartition1D, patchtable.h:1172-1201). The centres depend only on the sorted values at
// ranks (n-1)*i/nslices. They are found by an exact radix select over every dimension at once
// (four 8-bit digit passes, each a histogram and a resolve launch), or, for a dimension with
// more slices than the select keeps in LDS, by a full radix sort and a gather.
#include "pt_device.h"
#include "pt_kernels.h"

#include <hipcub/hipcub.hpp>

namespace {

constexpr int MM_THREADS = 256;

// Partial min/max of every dimension over the valid descriptor region [7,h) x [7,w), allowed
// patches only (TABLE_FOR_ALLOWED_PATCHES, patchtable.h:1766) when a mask is given. Blocks take
// rows in turn, lanes take the columns of a row.
__global__ void k_minmax(const float *__restrict__ desc, size_t plane, int Wp, int h, int w, int G,
                         const int32_t *__restrict__ allowed, int aidx, float *__restrict__ partials) {
    __shared__ float smin[MM_THREADS], smax[MM_THREADS];
    const int ew = w - 7, eh = h - 7;
    for (int i = 0; i < G; i++) {
        const float *P = desc + (size_t)i * plane;
        float mn = (float)1e20, mx = (float)-1e20;
        for (int y = blockIdx.x; y < eh; y += gridDim.x) {
            const float *row = P + (size_t)(y + 7) * Wp + 7;
            const int32_t *arow = allowed ? allowed + (size_t)y * ew : nullptr;
            for (int x = threadIdx.x; x < ew; x += blockDim.x) {
                if (arow && arow[x] != aidx) continue;
                const float v = row[x];
                mn = mn < v ? mn : v;
                mx = mx > v ? mx : v;
            }
        }
        smin[threadIdx.x] = mn;
        smax[threadIdx.x] = mx;
        __syncthreads();
        for (int o = MM_THREADS / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                float a = smin[threadIdx.x], b = smin[threadIdx.x + o];
                smin[threadIdx.x] = a < b ? a : b;
                a = smax[threadIdx.x];
                b = smax[threadIdx.x + o];
                smax[threadIdx.x] = a > b ? a : b;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            partials[((size_t)blockIdx.x * G + i) * 2 + 0] = smin[0];
            partials[((size_t)blockIdx.x * G + i) * 2 + 1] = smax[0];
        }
        __syncthreads();
    }
}

// Values of one dimension over the patches of the partition grid (every `step`-th patch in x
// and y, patchtable.h:1357-1370); patches outside the mask get +inf so that they sort after
// every allowed value and the order statistics of the allowed ones come first.
__global__ void k_extract(const float *__restrict__ desc, int Wp, int nx, int ny, int step, int bew,
                          const int32_t *__restrict__ allowed, int aidx, float *__restrict__ out) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nx * ny) return;
    long yy = k / nx, xx = k - yy * nx;
    float v = desc[(size_t)(7 + yy * step) * Wp + 7 + xx * step];
    if (allowed && allowed[(size_t)yy * step * bew + xx * step] != aidx) v = INFINITY;
    out[k] = v;
}

// The sorted values at ranks (npart-1)*j/nslices (PatchPartition1D, patchtable.h:1194-1196).
__global__ void k_gather_ranks(const float *__restrict__ sorted, int64_t npart, int nslices, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j <= nslices) out[j] = sorted[(npart - 1) * (int64_t)j / nslices];
}

// ---- Radix select of the quantile ranks -------------------------------------------------------
// Keys order floats as the radix sort does: negative values by flipped bits, the rest with the
// sign bit set, and -0.0 taken as +0.0 (rocPRIM's float digits treat them alike; its sort is
// stable, so a rank inside a run of zeros holding both signs is the zero at that input position,
// which the select cannot see: such a dimension is reported and sorted instead).
constexpr int RS_BINS = 256, RS_CHUNK = 8192, RS_THREADS = 256;

struct RsDims {
    int ns[PT_MAX_DIMS];   // slices per dimension; 0: not selected here
};
struct RsState {
    uint32_t *prefix;   // [G][ld] digits of rank j's key found so far
    int64_t *res;       // [G][ld] rank within the keys sharing that prefix
    int32_t *srank;     // [G][ld] rank j's slot in this pass's prefix list
    uint32_t *plist;    // [G][PT_RS_SLOTS] distinct prefixes, ascending
    int32_t *nsl;       // [G] entries of plist
    uint32_t *hist;     // [G][PT_RS_SLOTS][RS_BINS] counts of the next digit per prefix slot
    uint32_t *zc;       // [G][2] counts of -0.0 and +0.0
    int ld;
};

__device__ __forceinline__ uint32_t rs_key(uint32_t u) {
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_rs_init(RsState S, RsDims D, int64_t npart) {
    const int d = blockIdx.x, ns = D.ns[d];
    if (ns <= 0) return;
    for (int j = threadIdx.x; j <= ns; j += blockDim.x) {
        const size_t q = (size_t)d * S.ld + j;
        S.prefix[q] = 0;
        S.res[q] = (npart - 1) * (int64_t)j / ns;
        S.srank[q] = 0;
    }
    for (int i = threadIdx.x; i < PT_RS_SLOTS * RS_BINS; i += blockDim.x) S.hist[(size_t)d * PT_RS_SLOTS * RS_BINS + i] = 0;
    if (threadIdx.x == 0) {
        S.nsl[d] = 1;
        S.plist[d * PT_RS_SLOTS] = 0;
        S.zc[2 * d] = 0;
        S.zc[2 * d + 1] = 0;
    }
}

// Pass p counts digit p (bits 31-8p .. 24-8p) of the keys whose higher digits equal one of the
// prefixes still selected, in LDS per block, then adds the block's counts to the global histogram.
// Values are read where k_extract reads them (masked patches: +inf).
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const float *__restrict__ desc, size_t plane, int Wp, int nx,
                                                        int n, int step, int bew, const int32_t *__restrict__ allowed,
                                                        int aidx, RsState S, RsDims D, int pass) {
    __shared__ uint32_t lh[PT_RS_SLOTS * RS_BINS];
    __shared__ uint32_t pl[PT_RS_SLOTS];
    __shared__ uint32_t lz[2];
    const int d = blockIdx.y;
    if (D.ns[d] <= 0) return;   // uniform over the block
    const int nsl = S.nsl[d];
    for (int i = threadIdx.x; i < nsl * RS_BINS; i += RS_THREADS) lh[i] = 0;
    for (int i = threadIdx.x; i < nsl; i += RS_THREADS) pl[i] = S.plist[d * PT_RS_SLOTS + i];
    if (threadIdx.x < 2) lz[threadIdx.x] = 0;
    __syncthreads();
    const float *P = desc + (size_t)d * plane;
    const int k0 = blockIdx.x * RS_CHUNK, k1 = min(n, k0 + RS_CHUNK);
    const int dshift = 24 - 8 * pass;
    uint32_t zn = 0, zp = 0;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += RS_THREADS) {
        const int yy = k / nx, xx = k - yy * nx;
        float v = P[(size_t)(7 + yy * step) * Wp + 7 + (size_t)xx * step];
        if (allowed && allowed[(size_t)yy * step * bew + (size_t)xx * step] != aidx) v = INFINITY;
        const uint32_t u = __float_as_uint(v);
        zn += u == 0x80000000u;
        zp += u == 0u;
        const uint32_t key = rs_key(u);
        int slot = 0;
        if (pass > 0) {
            const uint32_t pre = key >> (32 - 8 * pass);
            int lo = 0, hi = nsl - 1;
            slot = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const uint32_t pv = pl[mid];
                if (pv == pre) {
                    slot = mid;
                    break;
                }
                if (pv < pre) lo = mid + 1;
                else hi = mid - 1;
            }
            if (slot < 0) continue;
        }
        atomicAdd(&lh[slot * RS_BINS + ((key >> dshift) & 255u)], 1u);
    }
    if (pass == 0) {
        if (zn) atomicAdd(&lz[0], zn);
        if (zp) atomicAdd(&lz[1], zp);
    }
    __syncthreads();
    uint32_t *gh = S.hist + (size_t)d * PT_RS_SLOTS * RS_BINS;
    for (int i = threadIdx.x; i < nsl * RS_BINS; i += RS_THREADS)
        if (lh[i]) atomicAdd(&gh[i], lh[i]);
    if (pass == 0 && threadIdx.x < 2 && lz[threadIdx.x]) atomicAdd(&S.zc[2 * d + threadIdx.x], lz[threadIdx.x]);
}

// Appends each rank's digit of this pass, clears the histogram, and lists the distinct prefixes
// for the next pass (ranks ascend, so their keys and prefixes do too). After the last pass the
// keys are complete: out[d][j] is the value, ambig[d] = 1 where a rank's key is zero and the
// dimension holds zeros of both signs.
__global__ void k_rs_resolve(RsState S, RsDims D, int pass, float *__restrict__ out, int *__restrict__ ambig) {
    const int d = blockIdx.x, ns = D.ns[d];
    if (ns <= 0) return;
    uint32_t *gh = S.hist + (size_t)d * PT_RS_SLOTS * RS_BINS;
    for (int j = threadIdx.x; j <= ns; j += blockDim.x) {
        const size_t q = (size_t)d * S.ld + j;
        const uint32_t *hh = gh + (size_t)S.srank[q] * RS_BINS;
        int64_t r = S.res[q];
        int b = 0;
        for (; b < RS_BINS - 1; b++) {
            if (r < (int64_t)hh[b]) break;
            r -= hh[b];
        }
        S.res[q] = r;
        S.prefix[q] = (S.prefix[q] << 8) | (uint32_t)b;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PT_RS_SLOTS * RS_BINS; i += blockDim.x) gh[i] = 0;
    if (threadIdx.x != 0) return;
    int m = 0;
    uint32_t last = 0;
    for (int j = 0; j <= ns; j++) {
        const size_t q = (size_t)d * S.ld + j;
        const uint32_t pre = S.prefix[q];
        if (m == 0 || pre != last) {
            S.plist[d * PT_RS_SLOTS + m] = pre;
            m++;
            last = pre;
        }
        S.srank[q] = m - 1;
    }
    S.nsl[d] = m;
    if (pass != 3) return;
    int amb = 0;
    for (int j = 0; j <= ns; j++) {
        const uint32_t key = S.prefix[(size_t)d * S.ld + j];
        uint32_t u = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
        if (u == 0u) {
            const uint32_t zn = S.zc[2 * d], zp = S.zc[2 * d + 1];
            if (zn && zp) amb = 1;
            else if (zn) u = 0x80000000u;
        }
        out[(size_t)d * S.ld + j] = __uint_as_float(u);
    }
    ambig[d] = amb;
}

}  // namespace

int pt_quantile_select(const float *desc, size_t plane, int Wp, int h, int w, int step, int G, const int *nslices,
                       const int32_t *allowed, int aidx, int64_t npart, int ld, float *values_dev, int *ambig_dev,
                       PtScratch &sc, hipStream_t s) {
    if (G > PT_MAX_DIMS) return pt_fail(PT_E_UNSUPPORTED, "too many dims");
    RsDims D{};
    bool any = false;
    for (int d = 0; d < G; d++) {
        if (nslices[d] + 1 > PT_RS_SLOTS || nslices[d] + 1 > ld)
            return pt_failf(PT_E_INVALID, "quantile select: %d slices in dimension %d", nslices[d], d);
        D.ns[d] = nslices[d];
        any |= nslices[d] > 0;
    }
    if (!any) return PT_OK;
    const int nx = (w - 7 + step - 1) / step, ny = (h - 7 + step - 1) / step;
    const long nl = (long)nx * ny;
    if (nl <= 0 || nl > 0x7fffffffL) return pt_fail(PT_E_UNSUPPORTED, "quantile select: grid size out of range");
    const int n = (int)nl;
    const size_t nq = (size_t)G * ld, hist_n = (size_t)G * PT_RS_SLOTS * RS_BINS;
    const size_t bytes = nq * (4 + 8 + 4) + (size_t)G * PT_RS_SLOTS * 4 + (size_t)G * 4 + hist_n * 4 + (size_t)G * 8 + 64;
    char *base = (char *)sc.get(13, bytes);
    if (!base) return pt_fail(PT_E_NOMEM, "quantile select: scratch allocation failed");
    RsState S;
    S.ld = ld;
    S.res = (int64_t *)base;   // 8-byte entries first
    S.prefix = (uint32_t *)(S.res + nq);
    S.srank = (int32_t *)(S.prefix + nq);
    S.plist = (uint32_t *)(S.srank + nq);
    S.nsl = (int32_t *)(S.plist + (size_t)G * PT_RS_SLOTS);
    S.zc = (uint32_t *)(S.nsl + G);
    S.hist = S.zc + (size_t)G * 2;
    hipLaunchKernelGGL(k_rs_init, dim3(G), dim3(256), 0, s, S, D, (int64_t)npart);
    const dim3 hgrid((n + RS_CHUNK - 1) / RS_CHUNK, G);
    for (int pass = 0; pass < 4; pass++) {
        hipLaunchKernelGGL(k_rs_hist, hgrid, dim3(RS_THREADS), 0, s, desc, plane, Wp, nx, n, step, w - 7, allowed, aidx,
                           S, D, pass);
        hipLaunchKernelGGL(k_rs_resolve, dim3(G), dim3(256), 0, s, S, D, pass, values_dev, ambig_dev);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_minmax_launch(const float *desc, size_t plane, int Wp, int h, int w, int G, const int32_t *allowed, int aidx,
                     float *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax, dim3(nblocks), dim3(MM_THREADS), 0, s, desc, plane, Wp, h, w, G, allowed, aidx,
                       partials);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_quantile_values(const float *desc, size_t plane, int Wp, int h, int w, int step, int dim, const int32_t *allowed,
                       int aidx, int64_t npart, int nslices, float *values_dev, PtScratch &sc, hipStream_t s) {
    int nx = (w - 7 + step - 1) / step, ny = (h - 7 + step - 1) / step;
    long n = (long)nx * ny;
    float *vals = (float *)sc.get(10, (size_t)n * 4);
    float *sorted = (float *)sc.get(11, (size_t)n * 4);
    if (!vals || !sorted) return pt_fail(PT_E_NOMEM, "quantiser: scratch allocation failed");
    hipLaunchKernelGGL(k_extract, dim3(pt_blocks(n, 256)), dim3(256), 0, s, desc + (size_t)dim * plane, Wp, nx, ny, step,
                       w - 7, allowed, aidx, vals);
    size_t tmp_bytes = 0;
    PT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, vals, sorted, (int)n, 0, 32, s));
    void *tmp = sc.get(12, tmp_bytes);
    if (!tmp) return pt_fail(PT_E_NOMEM, "quantiser: sort scratch allocation failed");
    PT_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, vals, sorted, (int)n, 0, 32, s));
    hipLaunchKernelGGL(k_gather_ranks, dim3(pt_blocks(nslices + 1, 256)), dim3(256), 0, s, sorted, npart, nslices,
                       values_dev);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
