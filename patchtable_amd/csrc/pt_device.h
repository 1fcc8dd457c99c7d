// Device-side helpers shared by the HIP kernels of the PatchTable library (gfx950).
#ifndef PT_DEVICE_H
#define PT_DEVICE_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pt_internal.h"

#define PT_HIP_TRY(expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return pt_failf(PT_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

#define PT_LAUNCH_CHECK() PT_HIP_TRY(hipGetLastError())

static inline unsigned pt_blocks(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// Adds the number of lanes of the block with `pred` set to *cnt: one atomic per block (atomics
// to one address from every wave of a large grid serialise in L2). Every thread of the block must
// call it (it synchronises the block); blockDim.x <= 1024.
__device__ __forceinline__ void pt_block_count_add(bool pred, int *cnt) {
    __shared__ int part[16];
    const unsigned long long b = __ballot(pred);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < (int)((blockDim.x + 63) / 64); i++) s += part[i];
        if (s) atomicAdd(cnt, s);
    }
}

// int(float) with the x86 cvttss2si result for out-of-range and NaN inputs (INT_MIN), which is
// what the reference's compiled get_bin computes (patchtable.h:1299).
__device__ __forceinline__ int pt_x86_ftoi(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int)f;
}

// get_bin (patchtable.h:1298-1303): FP32 (v-min)*16383/range truncated, clamped, remapped.
__device__ __forceinline__ int pt_get_bin(float v, float min_val, float range, const int32_t *__restrict__ to_bin) {
    int idx = pt_x86_ftoi((v - min_val) * (float)(PT_NMAP - 1) / range);
    if (idx < 0) idx = 0;
    else if (idx >= PT_NMAP) idx = PT_NMAP - 1;
    return to_bin[idx];
}

// Exact 8x8 RGB SSD with the lane order of patch_dist_exact (patchtable.h:330-373): lane j
// accumulates elements 4n+j (n = 0..5) of each row in order, lanes are summed last.
// a, b point at the patches' top-left pixels; pitches are row strides in floats (3*width).
__device__ __forceinline__ float pt_patch_ssd(const float *__restrict__ a, int apitch,
                                              const float *__restrict__ b, int bpitch) {
    float L0 = 0.f, L1 = 0.f, L2 = 0.f, L3 = 0.f;
#pragma unroll
    for (int dy = 0; dy < 8; dy++) {
        const float *ar = a + dy * apitch;
        const float *br = b + dy * bpitch;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            float d0 = ar[4 * n + 0] - br[4 * n + 0];
            float d1 = ar[4 * n + 1] - br[4 * n + 1];
            float d2 = ar[4 * n + 2] - br[4 * n + 2];
            float d3 = ar[4 * n + 3] - br[4 * n + 3];
            L0 = L0 + d0 * d0;
            L1 = L1 + d1 * d1;
            L2 = L2 + d2 * d2;
            L3 = L3 + d3 * d3;
        }
    }
    return ((L0 + L1) + L2) + L3;
}

// Same sum, abandoned once a partial sum reaches `bound`. Partial lane sums never exceed the
// final ones (adding non-negative floats is monotone), so a result >= bound means the exact SSD
// is >= bound too; a result < bound is the exact SSD.
__device__ __forceinline__ float pt_patch_ssd_bounded(const float *__restrict__ a, int apitch,
                                                      const float *__restrict__ b, int bpitch, float bound) {
    float L0 = 0.f, L1 = 0.f, L2 = 0.f, L3 = 0.f;
#pragma unroll
    for (int dy = 0; dy < 8; dy++) {
        const float *ar = a + dy * apitch;
        const float *br = b + dy * bpitch;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            float d0 = ar[4 * n + 0] - br[4 * n + 0];
            float d1 = ar[4 * n + 1] - br[4 * n + 1];
            float d2 = ar[4 * n + 2] - br[4 * n + 2];
            float d3 = ar[4 * n + 3] - br[4 * n + 3];
            L0 = L0 + d0 * d0;
            L1 = L1 + d1 * d1;
            L2 = L2 + d2 * d2;
            L3 = L3 + d3 * d3;
        }
        if ((dy & 1) && dy < 7) {
            float part = ((L0 + L1) + L2) + L3;
            if (part >= bound) return part;
        }
    }
    return ((L0 + L1) + L2) + L3;
}

// patch_dist_approx (patchtable.h:376-437) with the SSE lane order of descriptor mode: the first
// four dimensions as one 4-lane (a-b)^2; for C = 8, 12, 16, 20 the other 4-wide chunks are added
// lane-wise and the lanes summed last, otherwise the lanes are summed and the remaining
// dimensions added one by one (C >= 4).
__device__ __forceinline__ float pt_desc_dist(const float *a, const float *__restrict__ b, int C) {
    float L0, L1, L2, L3;
    {
        float e0 = a[0] - b[0], e1 = a[1] - b[1], e2 = a[2] - b[2], e3 = a[3] - b[3];
        L0 = e0 * e0;
        L1 = e1 * e1;
        L2 = e2 * e2;
        L3 = e3 * e3;
    }
    if (C == 8 || C == 12 || C == 16 || C == 20) {
        for (int n = 4; n < C; n += 4) {
            float e0 = a[n] - b[n], e1 = a[n + 1] - b[n + 1], e2 = a[n + 2] - b[n + 2], e3 = a[n + 3] - b[n + 3];
            L0 = L0 + e0 * e0;
            L1 = L1 + e1 * e1;
            L2 = L2 + e2 * e2;
            L3 = L3 + e3 * e3;
        }
        return ((L0 + L1) + L2) + L3;
    }
    float d = ((L0 + L1) + L2) + L3;
    for (int i = 4; i < C; i++) {
        float e = a[i] - b[i];
        d = d + e * e;
    }
    return d;
}

#endif
