// Gray-Code-Kernel (Walsh-Hadamard) projection on gfx950.
//
// Computes the same values, bit for bit, as gck_8<float,float,n,nchroma> (gck.h:29-1950): an
// 8x8 box filter by running sums along x then y, followed by one +-1 recurrence per further
// kernel in snake order. Each output value is a fixed sequence of FP32 adds/subtracts of its
// inputs, so the sequential recurrences are evaluated as independent chains (one lane per row
// or column chain) in exactly the reference's operand order; only the evaluation order across
// chains differs, which does not change any value.
//
// Layout: the descriptor is stored planar, plane c = channel c of the reference's HWC wh
// array, each plane (h+7) x (w+7) floats, row-major.
#include "pt_device.h"
#include "pt_kernels.h"

#include <vector>

namespace {

// RGB (or YUV) HWC -> three planes; RGB is converted with the double-precision weights of
// ARRAY_RGB2YUV (solver/array.h:1116-1127).
__global__ void k_to_planes(const float *__restrict__ img, long npx, int convert, float *__restrict__ planes) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float R = img[i * 3 + 0], G = img[i * 3 + 1], B = img[i * 3 + 2];
    float Y, U, V;
    if (convert) {
        double r = R, g = G, b = B;
        Y = (float)(0.299 * r + 0.587 * g + 0.114 * b);
        U = (float)(-0.14713 * r + -0.28886 * g + 0.436 * b);
        V = (float)(0.615 * r + -0.51499 * g + -0.10001 * b);
    } else {
        Y = R;
        U = G;
        V = B;
    }
    planes[i] = Y;
    planes[npx + i] = U;
    planes[2 * npx + i] = V;
}

// Steps of a running sum whose inputs a lane loads before adding them up: the chains are
// sequential, so loads in flight per lane are what hides the memory latency.
constexpr int GCK_BATCH = 32;
// ... along x, where a lane's steps are `delta` apart: 8 (32 spans a quarter row per batch and
// measured slower)
constexpr int GCK_BATCH_X = 8;

// Horizontal running box sum (gck.h:54-165), one lane per (row, channel) chain:
//   t(y,0) = in(y-3,0);  t(y,x) = (t(y,x-1) + in(y-3,x)) - in(y-3,x-8);
// reads outside the image are get_zero(), i.e. in(0,0,0) * 0.
__global__ void k_hpass(const float *__restrict__ yuv, int h, int w, float *__restrict__ temp, int Hp, int Wp) {
    int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= Hp * 3) return;
    int c = chain / Hp, y = chain - c * Hp;
    const float zero = yuv[0] * 0.0f;
    int ys = y - 3;
    float *out = temp + (size_t)c * Hp * Wp + (size_t)y * Wp;
    if (ys < 0 || ys >= h) {
        float t = zero;
        out[0] = t;
        for (int x = 1; x < Wp; x++) {
            t = (t + zero) - zero;
            out[x] = t;
        }
        return;
    }
    const float *row = yuv + (size_t)c * h * w + (size_t)ys * w;
    float t = row[0];
    out[0] = t;
    int x = 1;
    // x in [1, 8): the subtracted term is outside the image.
    for (; x < 8 && x < Wp; x++) {
        float add = x < w ? row[x] : zero;
        t = (t + add) - zero;
        out[x] = t;
    }
    // Interior: both terms inside the image (the subtracted one is in the lines just read).
    const int xend = w < Wp ? w : Wp;
    for (; x + GCK_BATCH <= xend; x += GCK_BATCH) {
        float add[GCK_BATCH], sub[GCK_BATCH];
#pragma unroll
        for (int k = 0; k < GCK_BATCH; k++) {
            add[k] = row[x + k];
            sub[k] = row[x + k - 8];
        }
#pragma unroll
        for (int k = 0; k < GCK_BATCH; k++) {
            t = (t + add[k]) - sub[k];
            out[x + k] = t;
        }
    }
    for (; x < Wp; x++) {
        float add = x < w ? row[x] : zero;
        float sub = (x - 8) < w ? row[x - 8] : zero;
        t = (t + add) - sub;
        out[x] = t;
    }
}

// Vertical running sum into descriptor channels 0..2 (gck.h:166-216), one lane per
// (channel, column) chain: o(0,x) = t(3,x); o(y,x) = (o(y-1,x) + t(y+3,x)) - t(y-5,x).
__global__ void k_vpass(const float *__restrict__ temp, int Hp, int Wp, float *__restrict__ desc) {
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= Wp * 3) return;
    int c = lane / Wp, x = lane - c * Wp;
    const float zero = temp[0] * 0.0f;
    const size_t plane = (size_t)Hp * Wp;
    const float *T = temp + c * plane + x;
    float *O = desc + c * plane + x;
    float o = T[(size_t)3 * Wp];
    O[0] = o;
    int y = 1;
    for (; y < 5 && y < Hp; y++) {
        float add = (y + 3 < Hp) ? T[(size_t)(y + 3) * Wp] : zero;
        o = (o + add) - zero;
        O[(size_t)y * Wp] = o;
    }
    // The term subtracted at y, t(y-5), was added at y - 8: after the first batch it comes from
    // registers.
    float prev[8];
    for (int first = 1; y + GCK_BATCH <= Hp - 3; y += GCK_BATCH, first = 0) {
        float add[GCK_BATCH], sub[GCK_BATCH];
#pragma unroll
        for (int k = 0; k < GCK_BATCH; k++) add[k] = T[(size_t)(y + k + 3) * Wp];
#pragma unroll
        for (int k = 0; k < GCK_BATCH; k++) {
            if (k < 8) sub[k] = first ? T[(size_t)(y + k - 5) * Wp] : prev[k];
            else sub[k] = add[k - 8];
        }
#pragma unroll
        for (int k = 0; k < GCK_BATCH; k++) {
            o = (o + add[k]) - sub[k];
            O[(size_t)(y + k) * Wp] = o;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) prev[k] = add[GCK_BATCH - 8 + k];
    }
    for (; y < Hp; y++) {
        float add = (y + 3 < Hp) ? T[(size_t)(y + 3) * Wp] : zero;
        float sub = T[(size_t)(y - 5) * Wp];
        o = (o + add) - sub;
        O[(size_t)y * Wp] = o;
    }
}

// Kernel k from kernel k-1 (gck.h:222-1946): k(p) = ((s*k(p-d)) + k1(p)) + s*k1(p-d), with
// p-d the position `delta` back along x; positions before the row start read get_zero().
// One lane per (row, residue x mod delta) chain; `cc` channels share the chain.
__global__ void k_recur_x(float *__restrict__ desc, int Hp, int Wp, int chk, int chp, int cc, int delta, int neg) {
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= Hp * delta) return;
    int y = lane / delta, r = lane - y * delta;
    const float zero = desc[0] * 0.0f;
    const size_t plane = (size_t)Hp * Wp;
    for (int c = 0; c < cc; c++) {
        float *K = desc + (size_t)(chk + c) * plane + (size_t)y * Wp;
        const float *P = desc + (size_t)(chp + c) * plane + (size_t)y * Wp;
        float kprev = zero, pprev = zero;
        int x = r;
        // GCK_BATCH_X steps' inputs are loaded ahead: they do not depend on the chain
        for (; x + (GCK_BATCH_X - 1) * delta < Wp; x += GCK_BATCH_X * delta) {
            float cur[GCK_BATCH_X];
#pragma unroll
            for (int u = 0; u < GCK_BATCH_X; u++) cur[u] = P[x + u * delta];
#pragma unroll
            for (int u = 0; u < GCK_BATCH_X; u++) {
                const float s1 = neg ? -kprev : kprev;
                const float s2 = neg ? -pprev : pprev;
                const float v = (s1 + cur[u]) + s2;
                K[x + u * delta] = v;
                kprev = v;
                pprev = cur[u];
            }
        }
        for (; x < Wp; x += delta) {
            float cur = P[x];
            float s1 = neg ? -kprev : kprev;
            float s2 = neg ? -pprev : pprev;
            float v = (s1 + cur) + s2;
            K[x] = v;
            kprev = v;
            pprev = cur;
        }
    }
}

// Same recurrence along y: one lane per (column, residue y mod delta) chain.
__global__ void k_recur_y(float *__restrict__ desc, int Hp, int Wp, int chk, int chp, int cc, int delta, int neg) {
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= Wp * delta) return;
    int r = lane / Wp, x = lane - r * Wp;
    const float zero = desc[0] * 0.0f;
    const size_t plane = (size_t)Hp * Wp;
    for (int c = 0; c < cc; c++) {
        float *K = desc + (size_t)(chk + c) * plane + x;
        const float *P = desc + (size_t)(chp + c) * plane + x;
        float kprev = zero, pprev = zero;
        int y = r;
        for (; y + (GCK_BATCH - 1) * delta < Hp; y += GCK_BATCH * delta) {
            float cur[GCK_BATCH];
#pragma unroll
            for (int u = 0; u < GCK_BATCH; u++) cur[u] = P[(size_t)(y + u * delta) * Wp];
#pragma unroll
            for (int u = 0; u < GCK_BATCH; u++) {
                const float s1 = neg ? -kprev : kprev;
                const float s2 = neg ? -pprev : pprev;
                const float v = (s1 + cur[u]) + s2;
                K[(size_t)(y + u * delta) * Wp] = v;
                kprev = v;
                pprev = cur[u];
            }
        }
        for (; y < Hp; y += delta) {
            float cur = P[(size_t)y * Wp];
            float s1 = neg ? -kprev : kprev;
            float s2 = neg ? -pprev : pprev;
            float v = (s1 + cur) + s2;
            K[(size_t)y * Wp] = v;
            kprev = v;
            pprev = cur;
        }
    }
}

// Walsh-Hadamard basis in Gray-code order with its alpha sign vectors and the snake order of
// the 2-D kernels (the construction of gck.py:37-67), giving for each kernel k >= 1 the axis,
// offset and sign of the recurrence that derives it from kernel k-1 (gck.py:389-400).
struct GckStep { int axis, delta, sign; };

void wh_alpha(int level, int k, std::vector<int> &cur, std::vector<int> &sign_level,
              std::vector<std::vector<int>> &out, int len) {
    std::vector<int> a1 = cur, a2 = cur;
    a1.push_back(sign_level[level]);
    a2.push_back(-sign_level[level]);
    sign_level[level] = -sign_level[level];
    if (len * 2 == (1 << k)) {
        out.push_back(a1);
        out.push_back(a2);
    } else {
        wh_alpha(level + 1, k, a1, sign_level, out, len * 2);
        wh_alpha(level + 1, k, a2, sign_level, out, len * 2);
    }
}

std::vector<GckStep> make_gck_steps() {
    std::vector<GckStep> steps;
    const int w = 8, k = 3;
    std::vector<std::vector<int>> alphas;
    std::vector<int> cur, sign_level(k, 1);
    wh_alpha(0, k, cur, sign_level, alphas, 1);
    std::vector<std::pair<int, int>> order;
    for (int r = 0; r < w; r++)
        for (int c = 0; c < w; c++) order.push_back({r, c});
    auto key = [&](const std::pair<int, int> &rc) {
        int m = rc.first > rc.second ? rc.first : rc.second;
        return m + 0.5 + (rc.first - rc.second + w) * (1.0 / (4 * w)) * (m % 2 == 0 ? -1 : 1);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](const std::pair<int, int> &a, const std::pair<int, int> &b) { return key(a) < key(b); });
    steps.push_back({-1, 0, 0});
    for (size_t i = 1; i < order.size(); i++) {
        bool horiz = order[i].second != order[i - 1].second;
        const std::vector<int> &a1 = alphas[horiz ? order[i - 1].second : order[i - 1].first];
        const std::vector<int> &a2 = alphas[horiz ? order[i].second : order[i].first];
        int idx = 0;
        while (a1[idx] == a2[idx]) idx++;
        steps.push_back({horiz ? 0 : 1, 1 << idx, a2[idx]});
    }
    return steps;
}

// Built once (thread-safe static initialisation): tables may be built from several threads.
const std::vector<GckStep> &gck_steps() {
    static const std::vector<GckStep> steps = make_gck_steps();
    return steps;
}

inline int gck_channel(int k, int nchroma) { return k < nchroma ? 3 * k : 3 * nchroma + (k - nchroma); }

}  // namespace

int pt_gck_supported(int n, int nchroma) {
    // The n, nchroma combinations gck() dispatches (gck.h:1957-2040, generated from gck.py:22-27).
    static const int dims[] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 20, 25, 30};
    if (nchroma < 0 || nchroma > 3) return 0;
    for (int d : dims)
        if (n == d - 2 * nchroma) return 1;
    return 0;
}

int pt_gck_launch(const float *img, int h, int w, int n, int nchroma, int convert_yuv, float *yuv_planes,
                  float *temp, float *desc, hipStream_t s) {
    if (!pt_gck_supported(n, nchroma))
        return pt_failf(PT_E_UNSUPPORTED, "gck: n, nchroma unsupported: %d, %d", n, nchroma);
    const int Hp = h + 7, Wp = w + 7;
    const long npx = (long)h * w;
    const int TB = 256;
    hipLaunchKernelGGL(k_to_planes, dim3(pt_blocks(npx, TB)), dim3(TB), 0, s, img, npx, convert_yuv, yuv_planes);
    hipLaunchKernelGGL(k_hpass, dim3(pt_blocks((long)Hp * 3, 64)), dim3(64), 0, s, yuv_planes, h, w, temp, Hp, Wp);
    hipLaunchKernelGGL(k_vpass, dim3(pt_blocks((long)Wp * 3, 64)), dim3(64), 0, s, temp, Hp, Wp, desc);
    const std::vector<GckStep> &st = gck_steps();
    for (int k = 1; k < n; k++) {
        int cc = k < nchroma ? 3 : 1;
        int chk = gck_channel(k, nchroma), chp = gck_channel(k - 1, nchroma);
        const GckStep &g = st[k];
        if (g.axis == 0)
            hipLaunchKernelGGL(k_recur_x, dim3(pt_blocks((long)Hp * g.delta, 64)), dim3(64), 0, s, desc, Hp, Wp, chk,
                               chp, cc, g.delta, g.sign < 0 ? 1 : 0);
        else
            hipLaunchKernelGGL(k_recur_y, dim3(pt_blocks((long)Wp * g.delta, 64)), dim3(64), 0, s, desc, Hp, Wp, chk,
                               chp, cc, g.delta, g.sign < 0 ? 1 : 0);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}
