// Learned patch key on gfx950: a 192 -> 64 -> 8 MLP (ReLU) that maps an 8x8 RGB patch to an
// 8-dimensional key, used in place of the GCK descriptor to index a table. Weights are stored in
// BF16, activations accumulate in FP32. The patch vector is x[dy*24 + dx*3 + c] = img(y+dy,
// x+dx, c) - 0.5 (the reference's HWC patch order, centred).
//
//   forward over an image (k_key_forward): one wave owns 16 patches of a row; layer 1 is
//     [16 x 192] x [192 x 64] on MFMA (v_mfma_f32_16x16x16_bf16: 12 K-steps x 4 N-tiles) with the
//     W1 fragments held in registers across the rows the block walks; bias + ReLU; layer 2
//     (64 -> 8) on VALU from LDS; the keys are written as the 8 descriptor planes of the table
//     ((h+7) x (w+7), patch (x, y) at (y+7, x+7)).
//   training: triplets (anchor, positive, negative) of patch vectors in BF16; k_train_forward
//     (VALU, FP32 accumulate over BF16 weights), k_triplet_grad (margin loss on squared key
//     distances), k_grad_w1 / k_grad_small (batch-reduced weight gradients), k_adam (FP32 master
//     weights, BF16 copies for the forward passes).
#include "pt_device.h"
#include "pt_kernels.h"

#include <algorithm>

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int KI = PT_KEY_IN, KH = PT_KEY_HIDDEN, KO = PT_KEY_OUT;

// float -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// ---- inference over an image ------------------------------------------------------------------
// Block: 4 waves over 64 consecutive patches of a row (wave w: patches x0 + 16w .. + 15), looping
// over rows y = blockIdx.y, blockIdx.y + gridDim.y, ... Each wave holds its W1 B-fragments in
// registers (48 x 4 BF16 per lane), loaded once, so the weights are read once per wave rather
// than once per row.
constexpr int KEY_ROW_BLOCKS = 64;

__global__ __launch_bounds__(256) void k_key_forward(const float *__restrict__ img, int h, int w,
                                                    const unsigned short *__restrict__ w1,   // KH x KI bf16
                                                    const float *__restrict__ b1,            // KH
                                                    const unsigned short *__restrict__ w2,   // KO x KH bf16
                                                    const float *__restrict__ b2,            // KO
                                                    float *__restrict__ planes, size_t plane, int Wp) {
    __shared__ float sh[4][16][KH + 1];   // hidden activations per wave
    __shared__ float sw2[KO * KH];
    for (int i = threadIdx.x; i < KO * KH; i += 256) sw2[i] = bf2f(w2[i]);
    const int ew = w - 7, eh = h - 7;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x0 = blockIdx.x * 64 + wave * 16;
    // MFMA 16x16x16 operand layout: A lane holds row lane % 16, k = 4*(lane/16) .. +3 of a 16-wide
    // K block; B lane holds column lane % 16 and the same k; D lane holds rows 4*(lane/16) + j.
    const int prow = lane & 15, kq = lane >> 4;
    v4s bfrag[KI / 16][4];
#pragma unroll
    for (int kb = 0; kb < KI / 16; kb++)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            const unsigned short *wr = w1 + (size_t)(nb * 16 + prow) * KI + kb * 16 + 4 * kq;
            bfrag[kb][nb] = v4s{(short)wr[0], (short)wr[1], (short)wr[2], (short)wr[3]};
        }
    float bias[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) bias[nb] = b1[nb * 16 + prow];
    __syncthreads();
    const int px = x0 + prow;
    const bool xvalid = px < ew;
    for (int y = blockIdx.y; y < eh; y += gridDim.y) {
        const float *pbase = img + ((size_t)y * w + (xvalid ? px : 0)) * 3;
        v4f acc[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[nb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KI / 16; kb++) {
            const int k = kb * 16 + 4 * kq;   // 4 consecutive elements inside one patch row of 24
            const int dy = k / 24, r = k % 24;
            v4s a = v4s{0, 0, 0, 0};
            if (xvalid) {
                const float *src = pbase + (size_t)dy * w * 3 + r;
                a[0] = (short)f2bf(src[0] - 0.5f);
                a[1] = (short)f2bf(src[1] - 0.5f);
                a[2] = (short)f2bf(src[2] - 0.5f);
                a[3] = (short)f2bf(src[3] - 0.5f);
            }
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfrag[kb][nb], acc[nb], 0, 0, 0);
        }
        // bias + ReLU into LDS: lane holds patches 4*kq + j, hidden unit nb*16 + prow
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = acc[nb][j] + bias[nb];
                sh[wave][4 * kq + j][nb * 16 + prow] = v > 0.f ? v : 0.f;
            }
        __syncthreads();
        // layer 2: 16 patches x 8 outputs per wave, 2 per lane
        for (int t = lane; t < 16 * KO; t += 64) {
            const int pi = t / KO, o = t % KO;
            float s = b2[o];
            for (int n = 0; n < KH; n++) s += sw2[o * KH + n] * sh[wave][pi][n];
            const int xx = x0 + pi;
            if (xx < ew) planes[(size_t)o * plane + (size_t)(y + 7) * Wp + xx + 7] = s;
        }
        __syncthreads();
    }
}

// ---- training ---------------------------------------------------------------------------------
// Patch vectors of the triplets (BF16): X[(3*t + role) * KI + k], role 0 anchor, 1 positive,
// 2 negative. Gathered from images by k_extract_triplets.
// pos[3*t + role] = XY_TO_INT(x, y) patch position (role 0 in a, roles 1, 2 in b); triplet t goes
// to slot[t] of X (t itself when slot is NULL).
__global__ void k_extract_triplets(const float *__restrict__ a, const float *__restrict__ b, int w,
                                   const int32_t *__restrict__ pos, const int64_t *__restrict__ slot, int nt,
                                   unsigned short *__restrict__ X) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nt * 3 * KI) return;
    const long s = i / KI;
    const int k = (int)(i % KI);
    const int v = pos[s];
    const float *img = (s % 3) == 0 ? a : b;
    const int x = PT_INT_TO_X(v), y = PT_INT_TO_Y(v);
    const int dy = k / 24, r = k % 24;
    const long t = s / 3, role = s % 3;
    const long d = slot ? slot[t] : t;
    X[(d * 3 + role) * KI + k] = f2bf(img[((size_t)(y + dy) * w + x) * 3 + r] - 0.5f);
}

// Forward of n samples (VALU, FP32 accumulate): pre-activation Z (n x KH) and keys K (n x KO).
__global__ void k_train_forward(const unsigned short *__restrict__ X, int n, const unsigned short *__restrict__ w1,
                                const float *__restrict__ b1, const unsigned short *__restrict__ w2,
                                const float *__restrict__ b2, float *__restrict__ Z, float *__restrict__ K) {
    const int s = blockIdx.x;   // one block of KH threads per sample
    const int j = threadIdx.x;
    __shared__ float xs[KI];
    __shared__ float hs[KH];
    if (s >= n) return;
    for (int k = j; k < KI; k += blockDim.x) xs[k] = bf2f(X[(size_t)s * KI + k]);
    __syncthreads();
    float z = b1[j];
    for (int k = 0; k < KI; k++) z += bf2f(w1[j * KI + k]) * xs[k];
    Z[(size_t)s * KH + j] = z;
    hs[j] = z > 0.f ? z : 0.f;
    __syncthreads();
    if (j < KO) {
        float o = b2[j];
        for (int m = 0; m < KH; m++) o += bf2f(w2[j * KH + m]) * hs[m];
        K[(size_t)s * KO + j] = o;
    }
}

// Triplet margin loss on squared key distances, averaged over nt triplets:
// L = mean(max(0, |ka-kp|^2 - |ka-kn|^2 + margin)); dK = dL/dK.
__global__ void k_triplet_grad(const float *__restrict__ K, int nt, float margin, float *__restrict__ dK,
                               float *__restrict__ loss_sum) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float *ka = K + (size_t)(3 * t) * KO, *kp = ka + KO, *kn = kp + KO;
    float dap = 0.f, dan = 0.f;
    for (int o = 0; o < KO; o++) {
        float e = ka[o] - kp[o], f = ka[o] - kn[o];
        dap += e * e;
        dan += f * f;
    }
    const float l = dap - dan + margin;
    const float sc = l > 0.f ? 1.f / nt : 0.f;
    float *ga = dK + (size_t)(3 * t) * KO, *gp = ga + KO, *gn = gp + KO;
    for (int o = 0; o < KO; o++) {
        ga[o] = sc * 2.f * (kn[o] - kp[o]);
        gp[o] = sc * -2.f * (ka[o] - kp[o]);
        gn[o] = sc * 2.f * (ka[o] - kn[o]);
    }
    if (l > 0.f) atomicAdd(loss_sum, l / nt);
}

// dH = (W2^T dK) * (Z > 0) per sample.
__global__ void k_grad_hidden(const float *__restrict__ dK, const float *__restrict__ Z, int n,
                              const unsigned short *__restrict__ w2, float *__restrict__ dH) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * KH) return;
    const long s = i / KH;
    const int j = (int)(i % KH);
    float g = 0.f;
    if (Z[i] > 0.f)
        for (int o = 0; o < KO; o++) g += bf2f(w2[o * KH + j]) * dK[s * KO + o];
    dH[i] = g;
}

// dW1[j][k] = sum_s dH[s][j] * X[s][k]; one thread per weight, samples in chunks.
__global__ void k_grad_w1(const float *__restrict__ dH, const unsigned short *__restrict__ X, int n,
                          float *__restrict__ gW1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= KH * KI) return;
    const int j = i / KI, k = i % KI;
    float g = 0.f;
    for (int s = 0; s < n; s++) g += dH[(size_t)s * KH + j] * bf2f(X[(size_t)s * KI + k]);
    gW1[i] = g;
}

// db1, dW2, db2 (one thread per parameter).
__global__ void k_grad_small(const float *__restrict__ dH, const float *__restrict__ Z, const float *__restrict__ dK,
                             int n, float *__restrict__ gb1, float *__restrict__ gW2, float *__restrict__ gb2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < KH) {
        float g = 0.f;
        for (int s = 0; s < n; s++) g += dH[(size_t)s * KH + i];
        gb1[i] = g;
    } else if (i < KH + KO * KH) {
        const int q = i - KH, o = q / KH, j = q % KH;
        float g = 0.f;
        for (int s = 0; s < n; s++) {
            const float z = Z[(size_t)s * KH + j];
            g += dK[(size_t)s * KO + o] * (z > 0.f ? z : 0.f);
        }
        gW2[q] = g;
    } else if (i < KH + KO * KH + KO) {
        const int o = i - KH - KO * KH;
        float g = 0.f;
        for (int s = 0; s < n; s++) g += dK[(size_t)s * KO + o];
        gb2[o] = g;
    }
}

// Adam step on FP32 master parameters; refreshes the BF16 copy (when given).
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2, float wd,
                       unsigned short *__restrict__ pbf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pn = p[i] - upd;
    p[i] = pn;
    if (pbf) pbf[i] = f2bf(pn);
}

__global__ void k_to_bf16(const float *__restrict__ in, int n, unsigned short *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

}  // namespace

int pt_key_forward_launch(const float *img, int h, int w, const unsigned short *w1, const float *b1,
                          const unsigned short *w2, const float *b2, float *planes, hipStream_t s) {
    const int ew = w - 7, eh = h - 7;
    if (ew <= 0 || eh <= 0) return pt_fail(PT_E_INVALID, "key: image smaller than a patch");
    const size_t plane = (size_t)(h + 7) * (w + 7);
    hipLaunchKernelGGL(k_key_forward, dim3((ew + 63) / 64, std::min(eh, KEY_ROW_BLOCKS)), dim3(256), 0, s, img, h, w,
                       w1, b1, w2, b2, planes, plane, w + 7);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_key_extract_launch(const float *a, const float *b, int w, const int32_t *pos, const int64_t *slot, int nt,
                          unsigned short *X, hipStream_t s) {
    const long n = (long)nt * 3 * KI;
    if (n <= 0) return PT_OK;
    hipLaunchKernelGGL(k_extract_triplets, dim3(pt_blocks(n, 256)), dim3(256), 0, s, a, b, w, pos, slot, nt, X);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_key_train_step(const unsigned short *X, int nt, float margin, const KeyParamsDev &P, const KeyWorkDev &Wk,
                      float lr, float beta1, float beta2, float eps, int step, float wd, hipStream_t s) {
    const int n = 3 * nt;
    hipLaunchKernelGGL(k_train_forward, dim3(n), dim3(KH), 0, s, X, n, P.w1bf, P.b1, P.w2bf, P.b2, Wk.Z, Wk.K);
    hipLaunchKernelGGL(k_triplet_grad, dim3(pt_blocks(nt, 256)), dim3(256), 0, s, Wk.K, nt, margin, Wk.dK, Wk.loss);
    hipLaunchKernelGGL(k_grad_hidden, dim3(pt_blocks((long)n * KH, 256)), dim3(256), 0, s, Wk.dK, Wk.Z, n, P.w2bf,
                       Wk.dH);
    hipLaunchKernelGGL(k_grad_w1, dim3(pt_blocks(KH * KI, 128)), dim3(128), 0, s, Wk.dH, X, n, Wk.gW1);
    hipLaunchKernelGGL(k_grad_small, dim3(pt_blocks(KH + KO * KH + KO, 64)), dim3(64), 0, s, Wk.dH, Wk.Z, Wk.dK, n,
                       Wk.gb1, Wk.gW2, Wk.gb2);
    if (lr > 0.f) {
        const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
        hipLaunchKernelGGL(k_adam, dim3(pt_blocks(KH * KI, 256)), dim3(256), 0, s, P.w1, Wk.gW1, P.m + 0, P.v + 0,
                           KH * KI, lr, beta1, beta2, eps, c1, c2, wd, P.w1bf);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(KH), 0, s, P.b1, Wk.gb1, P.m + KH * KI, P.v + KH * KI, KH, lr, beta1,
                           beta2, eps, c1, c2, 0.f, (unsigned short *)nullptr);
        hipLaunchKernelGGL(k_adam, dim3(pt_blocks(KO * KH, 256)), dim3(256), 0, s, P.w2, Wk.gW2, P.m + KH * KI + KH,
                           P.v + KH * KI + KH, KO * KH, lr, beta1, beta2, eps, c1, c2, wd, P.w2bf);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(KO), 0, s, P.b2, Wk.gb2, P.m + KH * KI + KH + KO * KH,
                           P.v + KH * KI + KH + KO * KH, KO, lr, beta1, beta2, eps, c1, c2, 0.f,
                           (unsigned short *)nullptr);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_key_to_bf16_launch(const float *in, int n, unsigned short *out, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(pt_blocks(n, 256)), dim3(256), 0, s, in, n, out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
