// Quantiser reductions on gfx950: per-dimension min/max over the database patches
// (set_min_max, patchtable.h:1758-1778) and the order statistics that place the slice centres
// (PatchPartition1D, patchtable.h:1172-1201). The centres depend only on the sorted values at
// ranks (n-1)*i/nslices. They are found by an exact radix select over every dimension at once
// (four 8-bit digit passes, each a histogram and a resolve launch), or, for a dimension with
// more slices than the select keeps in LDS, by a full radix sort and a gather.
#include "pt_device.h"
#include "pt_kernels.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace {

constexpr int MM_THREADS = 256;

// Partial min/max of dimension blockIdx.y over the valid descriptor region [7,h) x [7,w),
// allowed patches only (TABLE_FOR_ALLOWED_PATCHES, patchtable.h:1766) when a mask is given.
// Blocks take rows in turn, lanes take the columns of a row.
__global__ void k_minmax(const float *__restrict__ desc, size_t plane, int Wp, int h, int w, int G,
                         const int32_t *__restrict__ allowed, int aidx, float *__restrict__ partials) {
    __shared__ float smin[MM_THREADS], smax[MM_THREADS];
    const int ew = w - 7, eh = h - 7, i = blockIdx.y;
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
//
// Pass p (0..3) counts digit p (bits 31-8p .. 24-8p) of the keys whose higher digits form one of
// the prefixes the ranks have reached (the pass's slots); its resolve appends each rank's digit.
// Slots are found through byte tables: level L maps (slot of the first L-1 digits, digit L-1) to
// the slot of the first L digits, so a key's slot takes at most three LDS reads.
constexpr int RS_BINS = 256, RS_ROWS = 4, RS_THREADS = 256;
constexpr int RS_LUT = PT_RS_SLOTS * RS_BINS;   // entries of one table level

struct RsDims {
    int ns[PT_MAX_DIMS];   // slices per dimension; 0: not selected here
};
struct RsState {
    uint32_t *prefix;   // [G][ld] digits of rank j's key found so far
    int64_t *res;       // [G][ld] rank within the keys sharing that prefix
    int32_t *srank;     // [G][ld] rank j's slot in the next pass
    int32_t *cnt;       // [G][4] slots of pass p (1 in pass 0)
    int8_t *lut;        // [G][3][RS_LUT] slot tables, levels 1..3; -1: no rank has that prefix
    uint32_t *hist;     // [G][PT_RS_SLOTS][RS_BINS] counts of the pass's digit per slot
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
        S.cnt[4 * d] = 1;
        S.zc[2 * d] = 0;
        S.zc[2 * d + 1] = 0;
    }
}

// Histogram of pass p over `rows` rows of the partition grid of dimension blockIdx.y. Counters
// are 16-bit halves of LDS words (a block counts fewer than 65536 values), added once per group
// of lanes that share a bin for the two most common bins of a wave. Values are read where
// k_extract reads them (masked patches: +inf), eight loads in flight per lane.
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const float *__restrict__ desc, size_t plane, int Wp, int nx,
                                                        int ny, int rows, int step, int bew,
                                                        const int32_t *__restrict__ allowed, int aidx, RsState S,
                                                        RsDims D, int pass) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ uint32_t lz[2];
    const int d = blockIdx.y;
    if (D.ns[d] <= 0) return;   // uniform over the block
    const int nslot = S.cnt[4 * d + pass];
    const int c1 = pass >= 2 ? S.cnt[4 * d + 1] : 0, c2 = pass >= 3 ? S.cnt[4 * d + 2] : 0;
    uint32_t *lh = lds;                                      // nslot * 128 words
    int8_t *l1 = (int8_t *)(lds + nslot * (RS_BINS / 2));    // 256 entries
    int8_t *l2 = l1 + RS_BINS, *l3 = l2 + c1 * RS_BINS;      // c1 * 256, c2 * 256 entries
    for (int i = threadIdx.x; i < nslot * (RS_BINS / 2); i += RS_THREADS) lh[i] = 0;
    const int8_t *gl = S.lut + (size_t)d * 3 * RS_LUT;
    if (pass >= 1)
        for (int i = threadIdx.x; i < RS_BINS; i += RS_THREADS) l1[i] = gl[i];
    for (int i = threadIdx.x; i < c1 * RS_BINS; i += RS_THREADS) l2[i] = gl[RS_LUT + i];
    for (int i = threadIdx.x; i < c2 * RS_BINS; i += RS_THREADS) l3[i] = gl[2 * RS_LUT + i];
    if (threadIdx.x < 2) lz[threadIdx.x] = 0;
    __syncthreads();
    const float *P = desc + (size_t)d * plane;
    const int y0 = blockIdx.x * rows, y1 = min(ny, y0 + rows);
    const int dshift = 24 - 8 * pass;
    uint32_t zn = 0, zp = 0;
    for (int yy = y0; yy < y1; yy++) {
        const float *row = P + (size_t)(7 + yy * step) * Wp + 7;
        const int32_t *arow = allowed ? allowed + (size_t)yy * step * bew : nullptr;
        for (int c0 = 0; c0 < nx; c0 += RS_THREADS * 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int xx = c0 + u * RS_THREADS + (int)threadIdx.x;
                v[u] = xx < nx ? row[(size_t)xx * step] : 0.f;
                if (xx < nx && arow && arow[(size_t)xx * step] != aidx) v[u] = INFINITY;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int xx = c0 + u * RS_THREADS + (int)threadIdx.x;
                const uint32_t uu = __float_as_uint(v[u]);
                const uint32_t key = rs_key(uu);
                int bin = -1;
                if (xx < nx) {
                    zn += uu == 0x80000000u;
                    zp += uu == 0u;
                    int slot = 0;
                    if (pass >= 1) slot = l1[key >> 24];
                    if (pass >= 2 && slot >= 0) slot = l2[slot * RS_BINS + ((key >> 16) & 255u)];
                    if (pass >= 3 && slot >= 0) slot = l3[slot * RS_BINS + ((key >> 8) & 255u)];
                    if (slot >= 0) bin = slot * RS_BINS + (int)((key >> dshift) & 255u);
                }
#pragma unroll
                for (int round = 0; round < 2; round++) {
                    const unsigned long long act = __ballot(bin >= 0);   // lanes still to add
                    if (act == 0) break;
                    const int lead = __shfl(bin, __builtin_ctzll(act));  // >= 0: that lane is in act
                    const unsigned long long m = __ballot(bin == lead);
                    if (bin == lead) {
                        if ((int)(threadIdx.x & 63) == __builtin_ctzll(m))
                            atomicAdd(&lh[lead >> 1], (uint32_t)__popcll(m) << (16 * (lead & 1)));
                        bin = -1;
                    }
                }
                if (bin >= 0) atomicAdd(&lh[bin >> 1], 1u << (16 * (bin & 1)));
            }
        }
    }
    if (pass == 0) {
        if (zn) atomicAdd(&lz[0], zn);
        if (zp) atomicAdd(&lz[1], zp);
    }
    __syncthreads();
    uint32_t *gh = S.hist + (size_t)d * PT_RS_SLOTS * RS_BINS;
    for (int i = threadIdx.x; i < nslot * (RS_BINS / 2); i += RS_THREADS) {
        const uint32_t wv = lh[i];
        if (wv & 0xffffu) atomicAdd(&gh[2 * i], wv & 0xffffu);
        if (wv >> 16) atomicAdd(&gh[2 * i + 1], wv >> 16);
    }
    if (pass == 0 && threadIdx.x < 2 && lz[threadIdx.x]) atomicAdd(&S.zc[2 * d + threadIdx.x], lz[threadIdx.x]);
}

// Appends each rank's digit of pass p, clears the histogram, and (p < 3) writes the slot table
// of level p+1: ranks ascend, so their keys and prefixes do too, and equal prefixes are
// neighbours. After the last pass the keys are complete: out[d][j] is the value, ambig[d] = 1
// where a rank's key is zero and the dimension holds zeros of both signs.
__global__ __launch_bounds__(RS_THREADS) void k_rs_resolve(RsState S, RsDims D, int pass, float *__restrict__ out,
                                                           int *__restrict__ ambig) {
    __shared__ __align__(16) uint32_t lh[PT_RS_SLOTS * RS_BINS];
    __shared__ uint32_t lp[PT_RS_SLOTS];   // the ranks' prefixes after this pass
    __shared__ int ls[PT_RS_SLOTS];        // ... and their slots in this pass
    const int d = blockIdx.x, ns = D.ns[d];
    if (ns <= 0) return;
    uint32_t *gh = S.hist + (size_t)d * PT_RS_SLOTS * RS_BINS;
    const int nslot = S.cnt[4 * d + pass];
    for (int i = threadIdx.x; i < nslot * RS_BINS; i += RS_THREADS) {
        lh[i] = gh[i];
        gh[i] = 0;   // clear for the next pass
    }
    int8_t *nl = pass < 3 ? S.lut + ((size_t)d * 3 + pass) * RS_LUT : nullptr;   // level pass+1
    if (nl)
        for (int i = threadIdx.x; i < nslot * RS_BINS; i += RS_THREADS) nl[i] = -1;
    __syncthreads();
    for (int j = threadIdx.x; j <= ns; j += RS_THREADS) {
        const size_t q = (size_t)d * S.ld + j;
        const int sl = S.srank[q];
        const uint32_t *hh = lh + (size_t)sl * RS_BINS;
        int64_t r = S.res[q];
        int b = RS_BINS - 1;
        for (int c = 0; c < RS_BINS; c += 8) {
            const uint4 x0 = *(const uint4 *)(hh + c), x1 = *(const uint4 *)(hh + c + 4);
            const uint32_t v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const int64_t sum = (int64_t)v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7];
            if (r < sum || c == RS_BINS - 8) {
                int k = 0;
                for (; k < 7; k++) {
                    if (r < (int64_t)v[k]) break;
                    r -= v[k];
                }
                b = c + k;
                break;
            }
            r -= sum;
        }
        S.res[q] = r;
        const uint32_t pre = (S.prefix[q] << 8) | (uint32_t)b;
        S.prefix[q] = pre;
        lp[j] = pre;
        ls[j] = sl;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int m = 0;
    uint32_t last = 0;
    for (int j = 0; j <= ns; j++) {
        const uint32_t pre = lp[j];
        if (m == 0 || pre != last) {
            if (nl) nl[ls[j] * RS_BINS + (pre & 255u)] = (int8_t)m;
            m++;
            last = pre;
        }
        S.srank[(size_t)d * S.ld + j] = m - 1;
    }
    if (pass < 3) {
        S.cnt[4 * d + pass + 1] = m;
        return;
    }
    int amb = 0;
    for (int j = 0; j <= ns; j++) {
        const uint32_t key = lp[j];
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

// The value->bin map of every dimension (PatchPartition1D, patchtable.h:1233-1251), one entry
// per lane. The reference walks the centres once while the coordinate grows with the entry
// index, so its position for entry i is the first centre c with !(coord > centres[c]), or the
// last centre if there is none; each lane finds it on its own. The arithmetic is the
// reference's: coord = min + (i + 0.5) * range / 16384 in double, mid = the float sum of the two
// centres, divided by 2.0 in double.
__global__ void k_to_bin(const float *__restrict__ cen, int ldc, const int *__restrict__ nsv,
                         const float *__restrict__ minv, const float *__restrict__ rangev, int32_t *__restrict__ to_bin) {
    const int d = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PT_NMAP) return;
    const int ns = nsv[d];
    const float *c = cen + (size_t)d * ldc;
    const double coord = (double)minv[d] + ((double)i + 0.5) * (double)rangev[d] / (double)PT_NMAP;
    int ic = ns - 1;
    for (int k = 0; k < ns; k++)
        if (!(coord > (double)c[k])) {
            ic = k;
            break;
        }
    int pc = ic - 1;
    if (pc < 0) pc = 0;
    int nc = ic;
    if (nc == 0) nc++;
    if (ns == 1) nc = pc = 0;
    const float sum = c[pc] + c[nc];
    const double mid = (double)sum / 2.0;
    to_bin[(size_t)d * PT_NMAP + i] = coord < mid ? pc : nc;
}

}  // namespace

int pt_to_bin_launch(const float *centers, int ldc, const int *nslices, const float *min_val, const float *range, int G,
                     int32_t *to_bin, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bin, dim3(PT_NMAP / 256, G), dim3(256), 0, s, centers, ldc, nslices, min_val, range, to_bin);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_quantile_select(const float *desc, size_t plane, int Wp, int h, int w, int step, int G, const int *nslices,
                       const int32_t *allowed, int aidx, int64_t npart, int ld, float *values_dev, int *ambig_dev,
                       PtScratch &sc, hipStream_t s) {
    if (G > PT_MAX_DIMS) return pt_fail(PT_E_UNSUPPORTED, "too many dims");
    RsDims D{};
    bool any = false;
    int maxr = 1;   // slots a pass can hold: one in pass 0, at most a dimension's ranks after
    for (int d = 0; d < G; d++) {
        if (nslices[d] + 1 > PT_RS_SLOTS || nslices[d] + 1 > ld)
            return pt_failf(PT_E_INVALID, "quantile select: %d slices in dimension %d", nslices[d], d);
        D.ns[d] = nslices[d];
        any |= nslices[d] > 0;
        maxr = std::max(maxr, nslices[d] + 1);
    }
    if (!any) return PT_OK;
    const int nx = (w - 7 + step - 1) / step, ny = (h - 7 + step - 1) / step;
    if (nx <= 0 || ny <= 0) return PT_OK;
    if (nx > PT_RS_MAX_WIDTH) return pt_fail(PT_E_UNSUPPORTED, "quantile select: partition grid too wide");
    const int rows = std::min(RS_ROWS, 65535 / nx);   // fewer than 65536 values per block
    const size_t nq = (size_t)G * ld, hist_n = (size_t)G * PT_RS_SLOTS * RS_BINS;
    const size_t bytes = nq * (8 + 4 + 4) + (size_t)G * 16 + hist_n * 4 + (size_t)G * 8 + (size_t)G * 3 * RS_LUT + 64;
    char *base = (char *)sc.get(13, bytes);
    if (!base) return pt_fail(PT_E_NOMEM, "quantile select: scratch allocation failed");
    RsState S;
    S.ld = ld;
    S.res = (int64_t *)base;   // 8-byte entries first
    S.prefix = (uint32_t *)(S.res + nq);
    S.srank = (int32_t *)(S.prefix + nq);
    S.cnt = S.srank + nq;
    S.hist = (uint32_t *)(S.cnt + (size_t)G * 4);
    S.zc = S.hist + hist_n;
    S.lut = (int8_t *)(S.zc + (size_t)G * 2);
    hipLaunchKernelGGL(k_rs_init, dim3(G), dim3(256), 0, s, S, D, (int64_t)npart);
    const dim3 hgrid((ny + rows - 1) / rows, G);
    for (int pass = 0; pass < 4; pass++) {
        const int slots = pass == 0 ? 1 : maxr;
        const size_t lds = (size_t)slots * RS_BINS * 2 + (pass >= 1 ? RS_BINS : 0) + (pass >= 2 ? maxr * RS_BINS : 0) +
                           (pass >= 3 ? maxr * RS_BINS : 0);
        hipLaunchKernelGGL(k_rs_hist, hgrid, dim3(RS_THREADS), lds, s, desc, plane, Wp, nx, ny, rows, step, w - 7,
                           allowed, aidx, S, D, pass);
        hipLaunchKernelGGL(k_rs_resolve, dim3(G), dim3(RS_THREADS), 0, s, S, D, pass, values_dev, ambig_dev);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_minmax_launch(const float *desc, size_t plane, int Wp, int h, int w, int G, const int32_t *allowed, int aidx,
                     float *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax, dim3(nblocks, G), dim3(MM_THREADS), 0, s, desc, plane, Wp, h, w, G, allowed, aidx,
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
    PT_HIP_TRY(rocprim::radix_sort_keys(nullptr, tmp_bytes, vals, sorted, (int)n, 0, 32, s));
    void *tmp = sc.get(12, tmp_bytes);
    if (!tmp) return pt_fail(PT_E_NOMEM, "quantiser: sort scratch allocation failed");
    PT_HIP_TRY(rocprim::radix_sort_keys(tmp, tmp_bytes, vals, sorted, (int)n, 0, 32, s));
    hipLaunchKernelGGL(k_gather_ranks, dim3(pt_blocks(nslices + 1, 256)), dim3(256), 0, s, sorted, npart, nslices,
                       values_dev);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
