// Learned patch key on gfx950: a 192 -> 64 -> 8 MLP (ReLU) that maps an 8x8 RGB patch to an
// 8-dimensional key, used in place of the GCK descriptor to index a table. Weights are stored in
// BF16, activations accumulate in FP32. The patch vector is x[dy*24 + dx*3 + c] = img(y+dy,
// x+dx, c) - 0.5 (the reference's HWC patch order, centred).
//
//   forward over an image (k_key_forward): both layers on BF16 MFMA per 16 patches of a row
//     (the hidden units rounded to BF16 for layer 2), weights in registers; the keys are written as the 8 descriptor
//     planes of the table ((h+7) x (w+7), patch (x, y) at (y+7, x+7)).
//   training: triplets (anchor, positive, negative) of patch vectors in BF16; k_train_forward
//     (VALU, FP32 accumulate over BF16 weights), k_triplet_grad (margin loss on squared key
//     distances), k_grad_hidden, k_grad_partial (split over sample chunks; dW1 on FP32 MFMA) +
//     k_grad_reduce (batch-summed weight gradients), k_adam (FP32 master weights, BF16 copies for
//     the forward passes).
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
// One wave computes the keys of 16 consecutive patches of a row at a time (work items strided
// over the grid's waves), with both layers on MFMA and nothing in LDS:
//   layer 1, transposed: H^T (64 hidden x 16 patches) = W1 (64 x 192) . X^T (192 x 16) as
//     4 row tiles x 6 K-steps of v_mfma_f32_16x16x32_bf16; the W1 A-fragments stay in registers
//     for the wave's lifetime; the patch B-fragments are 8 consecutive RGB values of one patch row
//     (8 divides 24) converted to BF16 (round to nearest even) as they are loaded.
//   bias + ReLU in registers: lane l holds hidden units nb*16 + 4*(l/16) + j of patch l % 16.
//   layer 2: K^T (8 x 16) = W2 (8 x 64) . bf16(relu(H^T)) as 2 K-steps of 32 on
//     v_mfma_f32_16x16x32_bf16: lane l's 8 B slots of K-step m are its own hidden units
//     32m + 16(i/4) + 4*(l/16) + i%4 (the layer-1 accumulators of tiles 2m and 2m+1), and A holds
//     W2 at the same units (rows 8..15 zero). Lane l < 32 then holds outputs 4*(l/16) + j of
//     patch l % 16.
// A and B operands are filled under one convention (lane group l/16 owns the same K indices in
// both), so the sum over K is independent of how the hardware numbers K inside a group.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float v8f __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void k_key_forward(const float *__restrict__ img, int h, int w,
                                                    const unsigned short *__restrict__ w1,   // KH x KI bf16
                                                    const float *__restrict__ b1,            // KH
                                                    const unsigned short *__restrict__ w2,   // KO x KH bf16
                                                    const float *__restrict__ b2,            // KO
                                                    float *__restrict__ planes, size_t plane, int Wp) {
    const int ew = w - 7, eh = h - 7;
    const int lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    bf8 afrag[4][KI / 32];
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int kb = 0; kb < KI / 32; kb++) {
            const unsigned short *wr = w1 + (size_t)(nb * 16 + col) * KI + kb * 32 + 8 * grp;
            bf8 v;
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = __builtin_bit_cast(__bf16, wr[i]);
            afrag[nb][kb] = v;
        }
    float bias1[4][4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int j = 0; j < 4; j++) bias1[nb][j] = b1[nb * 16 + 4 * grp + j];
    // layer-2 A fragments: W2 row col (zero for col >= 8) at the hidden units the B operand's
    // slots hold: K-step m, lane group grp, slot i -> unit 32m + 16(i >> 2) + 4grp + (i & 3)
    bf8 w2f[2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const unsigned short v = w2[(col & (KO - 1)) * KH + 32 * m + 16 * (i >> 2) + 4 * grp + (i & 3)];
            w2f[m][i] = __builtin_bit_cast(__bf16, col < KO ? v : (unsigned short)0);
        }
    float bias2[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float v = b2[(4 * grp + j) & (KO - 1)];
        bias2[j] = 4 * grp + j < KO ? v : 0.f;
    }

    const int ngx = (ew + 15) / 16;
    const long items = (long)ngx * eh;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long it = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); it < items; it += nwaves) {
        const int y = (int)(it / ngx), x0 = (int)(it % ngx) * 16;
        const int px = x0 + col;
        const bool valid = px < ew;
        // lanes past the row end read the last patch (in bounds) and skip the store
        const float *pbase = img + ((size_t)y * w + (valid ? px : ew - 1)) * 3;
        v4f acc[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) acc[nb] = v4f{bias1[nb][0], bias1[nb][1], bias1[nb][2], bias1[nb][3]};
#pragma unroll
        for (int kb = 0; kb < KI / 32; kb++) {
            const int k = kb * 32 + 8 * grp;   // 8 consecutive values inside one patch row of 24
            const float *src = pbase + (size_t)(k / 24) * w * 3 + k % 24;
            v8f xv;
#pragma unroll
            for (int i = 0; i < 8; i++) xv[i] = src[i] - 0.5f;
            const bf8 bv = __builtin_convertvector(xv, bf8);
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag[nb][kb], bv, acc[nb], 0, 0, 0);
        }
        // layer 2 on two BF16 K-steps of 32: the ReLU'd hidden units, rounded to BF16 (nearest
        // even), are the B operand straight from the layer-1 accumulators (slot i of K-step m is
        // acc[2m + (i >> 2)][i & 3])
        v4f out = v4f{bias2[0], bias2[1], bias2[2], bias2[3]};
#pragma unroll
        for (int m = 0; m < 2; m++) {
            v8f hv;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const float z = acc[2 * m + (i >> 2)][i & 3];
                hv[i] = z > 0.f ? z : 0.f;
            }
            out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2f[m], __builtin_convertvector(hv, bf8), out, 0, 0, 0);
        }
        if (grp < 2 && valid) {
            float *dst = planes + (size_t)(y + 7) * Wp + px + 7;
#pragma unroll
            for (int j = 0; j < 4; j++) dst[(size_t)(4 * grp + j) * plane] = out[j];
        }
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

// Weight gradients, split over chunks of KEY_CHUNK samples: block c writes the partial sums of its
// chunk to part[c * KEY_NGRAD ..] (dW1 64 x 192, then db1 64, dW2 8 x 64, db2 8), and
// k_grad_reduce adds the chunks in order (deterministic).
//   dW1 = dH^T . X on MFMA in FP32 (v_mfma_f32_16x16x4_f32): wave w owns hidden rows 16w..16w+15
//   and all 12 column tiles; a K-step takes 4 samples (A: dH, B: X widened from BF16 exactly).
//   db1, dW2, db2 on VALU, one output per thread.
constexpr int KEY_CHUNK = 128;
constexpr int KEY_NGRAD = KH * KI + KH + KO * KH + KO;

__global__ __launch_bounds__(256) void k_grad_partial(const float *__restrict__ dH, const unsigned short *__restrict__ X,
                                                     const float *__restrict__ Z, const float *__restrict__ dK, int n,
                                                     float *__restrict__ part) {
    const int c = blockIdx.x;
    const int s0 = c * KEY_CHUNK, s1 = min(n, s0 + KEY_CHUNK);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, grp = lane >> 4;
    float *P = part + (size_t)c * KEY_NGRAD;
    v4f acc[KI / 16];
#pragma unroll
    for (int kt = 0; kt < KI / 16; kt++) acc[kt] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int s = s0; s < s1; s += 4) {
        const int si = s + grp;
        const bool ok = si < s1;
        const float a = ok ? dH[(size_t)si * KH + wave * 16 + col] : 0.f;
#pragma unroll
        for (int kt = 0; kt < KI / 16; kt++) {
            const float b = ok ? bf2f(X[(size_t)si * KI + kt * 16 + col]) : 0.f;
            acc[kt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[kt], 0, 0, 0);
        }
    }
    // D: lane holds rows (hidden) wave*16 + 4*grp + j, column kt*16 + col
#pragma unroll
    for (int kt = 0; kt < KI / 16; kt++)
#pragma unroll
        for (int j = 0; j < 4; j++) P[(size_t)(wave * 16 + 4 * grp + j) * KI + kt * 16 + col] = acc[kt][j];
    for (int t = threadIdx.x; t < KH + KO * KH + KO; t += blockDim.x) {
        float g = 0.f;
        if (t < KH) {
            for (int s = s0; s < s1; s++) g += dH[(size_t)s * KH + t];
        } else if (t < KH + KO * KH) {
            const int q = t - KH, o = q / KH, j = q % KH;
            for (int s = s0; s < s1; s++) {
                const float z = Z[(size_t)s * KH + j];
                g += dK[(size_t)s * KO + o] * (z > 0.f ? z : 0.f);
            }
        } else {
            const int o = t - KH - KO * KH;
            for (int s = s0; s < s1; s++) g += dK[(size_t)s * KO + o];
        }
        P[KH * KI + t] = g;
    }
}

__global__ void k_grad_reduce(const float *__restrict__ part, int nchunks, float *__restrict__ gW1,
                              float *__restrict__ gb1, float *__restrict__ gW2, float *__restrict__ gb2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= KEY_NGRAD) return;
    float g = 0.f;
    for (int c = 0; c < nchunks; c++) g += part[(size_t)c * KEY_NGRAD + i];
    if (i < KH * KI) gW1[i] = g;
    else if (i < KH * KI + KH) gb1[i - KH * KI] = g;
    else if (i < KH * KI + KH + KO * KH) gW2[i - KH * KI - KH] = g;
    else gb2[i - KH * KI - KH - KO * KH] = g;
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
    // 16-patch work items strided over at most 2048 blocks of 4 waves
    const long items = (long)((ew + 15) / 16) * eh;
    const int blocks = (int)std::min<long>(2048, (items + 3) / 4);
    hipLaunchKernelGGL(k_key_forward, dim3(blocks), dim3(256), 0, s, img, h, w, w1, b1, w2, b2, planes, plane, w + 7);
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
    const int nchunks = pt_key_grad_chunks(n);
    hipLaunchKernelGGL(k_grad_partial, dim3(nchunks), dim3(256), 0, s, Wk.dH, X, Wk.Z, Wk.dK, n, Wk.part);
    hipLaunchKernelGGL(k_grad_reduce, dim3(pt_blocks(KEY_NGRAD, 256)), dim3(256), 0, s, Wk.part, nchunks, Wk.gW1,
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

int pt_key_grad_chunks(int nsamples) { return (nsamples + KEY_CHUNK - 1) / KEY_CHUNK; }
size_t pt_key_grad_part_floats(int nsamples) { return (size_t)pt_key_grad_chunks(nsamples) * KEY_NGRAD; }

int pt_key_to_bf16_launch(const float *in, int n, unsigned short *out, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(pt_blocks(n, 256)), dim3(256), 0, s, in, n, out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
