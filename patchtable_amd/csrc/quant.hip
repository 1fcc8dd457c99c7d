// Quantiser reductions on gfx950: per-dimension min/max over the database patches
// (set_min_max, patchtable.h:1758-1778) and the order statistics that place the slice centres
// (PatchPartition1D, patchtable.h:1172-1201). The centres depend only on the sorted values at
// ranks (n-1)*i/nslices, so each dimension is radix-sorted once and the needed ranks gathered.
#include "pt_device.h"
#include "pt_kernels.h"

#include <hipcub/hipcub.hpp>

namespace {

constexpr int MM_THREADS = 256;

// Partial min/max of one dimension over the valid descriptor region [7,h) x [7,w), allowed
// patches only (TABLE_FOR_ALLOWED_PATCHES, patchtable.h:1766) when a mask is given.
__global__ void k_minmax(const float *__restrict__ desc, size_t plane, int Wp, int h, int w, int G,
                         const int32_t *__restrict__ allowed, int aidx, float *__restrict__ partials) {
    __shared__ float smin[MM_THREADS], smax[MM_THREADS];
    const int ew = w - 7;
    const long n = (long)(h - 7) * ew;
    for (int i = 0; i < G; i++) {
        const float *P = desc + (size_t)i * plane;
        float mn = (float)1e20, mx = (float)-1e20;
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
            long y = k / ew, x = k - y * ew;
            if (allowed && allowed[k] != aidx) continue;
            float v = P[(size_t)(y + 7) * Wp + x + 7];
            mn = mn < v ? mn : v;
            mx = mx > v ? mx : v;
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

__global__ void k_gather(const float *__restrict__ sorted, const int64_t *__restrict__ ranks, int nr, float *__restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) out[r] = sorted[ranks[r]];
}

}  // namespace

int pt_minmax_launch(const float *desc, size_t plane, int Wp, int h, int w, int G, const int32_t *allowed, int aidx,
                     float *partials, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax, dim3(nblocks), dim3(MM_THREADS), 0, s, desc, plane, Wp, h, w, G, allowed, aidx,
                       partials);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_quantile_values(const float *desc, size_t plane, int Wp, int h, int w, int step, int dim, const int32_t *allowed,
                       int aidx, const int64_t *ranks_dev, int nranks, float *values_dev, PtScratch &sc, hipStream_t s) {
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
    hipLaunchKernelGGL(k_gather, dim3(pt_blocks(nranks, 256)), dim3(256), 0, s, sorted, ranks_dev, nranks, values_dev);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
