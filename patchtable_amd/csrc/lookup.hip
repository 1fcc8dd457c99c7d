// Query lookup on gfx950 (replaces lookup_templated<0,0>, patchtable.h:3468-4217).
//
// The reference visits query patches in raster order (then reverse raster order) and lets each
// patch push its best offset to its 3x3 neighbours before they are visited, so a good match
// propagates along a whole row within one pass. Here every query patch is processed in
// parallel and propagation runs in rounds over double-buffered fields:
//   candidates: the table cell's patch (patchtable.h:3664-3704), then the k-coherence list of
//     the best with the triangle early-out (patchtable.h:3748-3851), then the spatial window
//     around the best (patchtable.h:3853-3885);
//   rounds: jump flooding -- the 8 neighbours at distance J (J = 8, 4, 2, 1, then 1 ...) each
//     propose their offset; then the spatial window around the best again.
// Candidates replace the best only when strictly closer and every distance is the exact SSD
// with the reference's lane order, so NNF_DIST is always the exact SSD of the stored match.
#include "pt_device.h"
#include "pt_kernels.h"

#include <vector>

namespace {

__constant__ int c_dx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int c_dy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

// Spatial search (patchtable.h:3861-3885), including the reference's "else if" clamping.
__device__ __forceinline__ void spatial_search(const LookupArgs &L, const float *pa, int apitch, int &xb, int &yb,
                                               float &db) {
    const int bew = L.bw - 7, beh = L.bh - 7, r = L.spatial;
    int symin = yb - r, symax = yb + r + 1, sxmin = xb - r, sxmax = xb + r + 1;
    const int xb0 = xb, yb0 = yb;
    if (symin < 0) symin = 0;
    else if (symax > beh) symax = beh;
    if (sxmin < 0) sxmin = 0;
    else if (sxmax > bew) sxmax = bew;
    // The reference's "else if" leaves the far side unclamped when the window overhangs both
    // sides (a database fewer than 2r+1 patches tall or wide), where it reads out of bounds.
    if (symax > beh) symax = beh;
    if (sxmax > bew) sxmax = bew;
    const int bpitch = L.bw * 3;
    for (int ys = symin; ys < symax; ys++)
        for (int xs = sxmin; xs < sxmax; xs++) {
            if (xs == xb0 && ys == yb0) continue;
            float d = pt_patch_ssd_bounded(pa, apitch, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
            }
        }
}

__global__ void k_candidates(LookupArgs L, int32_t *__restrict__ nx, int32_t *__restrict__ ny, float *__restrict__ nd) {
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)ann_w * ann_h) return;
    const int x = (int)(i % ann_w), y = (int)(i / ann_w);
    const int apitch = L.aw * 3, bpitch = L.bw * 3, bew = L.bw - 7;
    const float *pa = L.a + ((size_t)y * L.aw + x) * 3;
    // table cell of the query descriptor at (y+7, x+7)
    const size_t plane = (size_t)(L.ah + 7) * (L.aw + 7);
    const size_t off = (size_t)(y + 7) * (L.aw + 7) + x + 7;
    int64_t cell = 0;
    for (int d = 0; d < L.q.G; d++)
        cell += (int64_t)pt_get_bin(L.a_desc[d * plane + off], L.q.min_val[d], L.q.range[d],
                                    L.q.to_bin + (size_t)d * PT_NMAP) *
                L.q.stride[d];
    int v = L.table[cell];
    int xb = PT_INT_TO_X(v) - 7, yb = PT_INT_TO_Y(v) - 7;
    float db = pt_patch_ssd(pa, apitch, L.b + ((size_t)yb * L.bw + xb) * 3, bpitch);
    if (L.K > 0 && L.kcoh_iter0) {
        const int32_t *kp = L.kpos + ((size_t)yb * bew + xb) * L.K;
        const float *kd = L.kdist + ((size_t)yb * bew + xb) * L.K;
        for (int k = 0; k < L.K; k++) {
            if ((double)kd[k] >= L.triangle_factor * (double)db) break;
            int xs = PT_INT_TO_X(kp[k]), ys = PT_INT_TO_Y(kp[k]);
            float d = pt_patch_ssd_bounded(pa, apitch, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
            }
        }
    }
    if (L.spatial) spatial_search(L, pa, apitch, xb, yb, db);
    nx[i] = xb;
    ny[i] = yb;
    nd[i] = db;
}

__global__ void k_propagate(LookupArgs L, int J, const int32_t *__restrict__ cx, const int32_t *__restrict__ cy,
                            const float *__restrict__ cd, int32_t *__restrict__ nx, int32_t *__restrict__ ny,
                            float *__restrict__ nd) {
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)ann_w * ann_h) return;
    const int x = (int)(i % ann_w), y = (int)(i / ann_w);
    const int apitch = L.aw * 3, bpitch = L.bw * 3, bew = L.bw - 7, beh = L.bh - 7;
    const float *pa = L.a + ((size_t)y * L.aw + x) * 3;
    int xb = cx[i], yb = cy[i];
    float db = cd[i];
    for (int t = 0; t < 8; t++) {
        int qx = x + J * c_dx[t], qy = y + J * c_dy[t];
        if (qx < 0 || qy < 0 || qx >= ann_w || qy >= ann_h) continue;
        long qi = (long)qy * ann_w + qx;
        int xs = cx[qi] - J * c_dx[t], ys = cy[qi] - J * c_dy[t];
        if (xs < 0 || ys < 0 || xs >= bew || ys >= beh) continue;
        if (xs == xb && ys == yb) continue;
        float d = pt_patch_ssd_bounded(pa, apitch, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
        if (d < db) {
            db = d;
            xb = xs;
            yb = ys;
        }
    }
    if (L.spatial) spatial_search(L, pa, apitch, xb, yb, db);
    nx[i] = xb;
    ny[i] = yb;
    nd[i] = db;
}

__global__ void k_write_ann(const int32_t *__restrict__ nx, const int32_t *__restrict__ ny, const float *__restrict__ nd,
                            long n, double *__restrict__ ann) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ann[i * 3 + 0] = (double)nx[i];
    ann[i * 3 + 1] = (double)ny[i];
    ann[i * 3 + 2] = (double)nd[i];
}

__global__ void k_ssd_pairs(const float *__restrict__ a, int aw, const float *__restrict__ b, int bw,
                            const int32_t *__restrict__ ax, const int32_t *__restrict__ ay, const int32_t *__restrict__ bx,
                            const int32_t *__restrict__ by, int n, float *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = pt_patch_ssd(a + ((size_t)ay[i] * aw + ax[i]) * 3, aw * 3, b + ((size_t)by[i] * bw + bx[i]) * 3, bw * 3);
}

}  // namespace

// Jump schedule of the propagation rounds: 8, 4, 2, 1 then unit jumps (fewer rounds start lower).
void pt_jump_schedule(int rounds, std::vector<int> &J) {
    J.clear();
    int big = rounds >= 4 ? 8 : (rounds == 3 ? 4 : (rounds == 2 ? 2 : 1));
    for (int j = big; j >= 1 && (int)J.size() < rounds; j >>= 1) J.push_back(j);
    while ((int)J.size() < rounds) J.push_back(1);
}

int pt_lookup_launch(const LookupArgs &L, int32_t *nx, int32_t *ny, float *nd, int32_t *nx2, int32_t *ny2, float *nd2,
                     double *ann_out, hipStream_t s, hipEvent_t ev_after_candidates, hipEvent_t ev_after_rounds) {
    const long n = (long)(L.aw - 7) * (L.ah - 7);
    const int TB = 64;
    hipLaunchKernelGGL(k_candidates, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, L, nx, ny, nd);
    if (ev_after_candidates) PT_HIP_TRY(hipEventRecord(ev_after_candidates, s));
    std::vector<int> J;
    pt_jump_schedule(L.rounds, J);
    int32_t *cx = nx, *cy = ny, *ox = nx2, *oy = ny2;
    float *cdist = nd, *odist = nd2;
    for (int j : J) {
        hipLaunchKernelGGL(k_propagate, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, L, j, cx, cy, cdist, ox, oy, odist);
        std::swap(cx, ox);
        std::swap(cy, oy);
        std::swap(cdist, odist);
    }
    if (ev_after_rounds) PT_HIP_TRY(hipEventRecord(ev_after_rounds, s));
    hipLaunchKernelGGL(k_write_ann, dim3(pt_blocks(n, 256)), dim3(256), 0, s, cx, cy, cdist, n, ann_out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_ssd_pairs_launch(const float *a, int aw, const float *b, int bw, const int32_t *ax, const int32_t *ay,
                        const int32_t *bx, const int32_t *by, int n, float *out, hipStream_t s) {
    if (n <= 0) return PT_OK;
    hipLaunchKernelGGL(k_ssd_pairs, dim3(pt_blocks(n, 256)), dim3(256), 0, s, a, aw, b, bw, ax, ay, bx, by, n, out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
