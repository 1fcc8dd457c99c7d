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
//
// Work layout: a workgroup of 256 lanes owns a 64 x 4 tile of query patches and stages the
// tile's 71 x 11 query pixels in LDS (one float4 per pixel), so each SSD reads only the
// database patch from global memory. The field is packed as XY_TO_INT position + distance, plus
// a flag recording that the spatial window around the stored best has already been searched
// without improvement; a later round whose best is unchanged skips that window (it would find
// nothing strictly closer), which leaves every result unchanged. A second flag records whether
// the best changed in the round that produced the field (used to skip repeated candidates).
#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_tile.h"

#include <vector>

namespace {

using namespace pt_tile;

__constant__ int c_dx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int c_dy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

// Spatial search (patchtable.h:3861-3885), including the reference's "else if" clamping; the
// far side is also clamped where the reference would read out of bounds (a database fewer than
// 2r+1 patches tall or wide). Returns true if the best changed.
__device__ __forceinline__ bool spatial_search(const LookupArgs &L, const float4 *aq, int &xb, int &yb, float &db) {
    const int bew = L.bw - 7, beh = L.bh - 7, r = L.spatial;
    int symin = yb - r, symax = yb + r + 1, sxmin = xb - r, sxmax = xb + r + 1;
    const int xb0 = xb, yb0 = yb;
    if (symin < 0) symin = 0;
    else if (symax > beh) symax = beh;
    if (sxmin < 0) sxmin = 0;
    else if (sxmax > bew) sxmax = bew;
    if (symax > beh) symax = beh;
    if (sxmax > bew) sxmax = bew;
    const int bpitch = L.bw * 3;
    bool changed = false;
    for (int ys = symin; ys < symax; ys++)
        for (int xs = sxmin; xs < sxmax; xs++) {
            if (xs == xb0 && ys == yb0) continue;
            float d = ssd_staged<true>(aq, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
                changed = true;
            }
        }
    return changed;
}

__global__ __launch_bounds__(THREADS) void k_candidates(LookupArgs L, NnfField out) {
    __shared__ float4 tile[AX * AY];
    stage_tile(L.a, L.ah, L.aw, tile);
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    const int lx = threadIdx.x % TILE_X, ly = threadIdx.x / TILE_X;
    const int x = blockIdx.x * TILE_X + lx, y = blockIdx.y * TILE_Y + ly;
    if (x >= ann_w || y >= ann_h) return;
    const float4 *aq = tile + ly * AX + lx;
    const int bpitch = L.bw * 3, bew = L.bw - 7;
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
    float db = ssd_staged<false>(aq, L.b + ((size_t)yb * L.bw + xb) * 3, bpitch, 0.f);
    if (L.K > 0 && L.kcoh_iter0) {
        const int32_t *kp = L.kpos + ((size_t)yb * bew + xb) * L.K;
        const float *kd = L.kdist + ((size_t)yb * bew + xb) * L.K;
        for (int k = 0; k < L.K; k++) {
            if ((double)kd[k] >= L.triangle_factor * (double)db) break;
            int xs = PT_INT_TO_X(kp[k]), ys = PT_INT_TO_Y(kp[k]);
            float d = ssd_staged<true>(aq, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
            }
        }
    }
    bool done = false;
    if (L.spatial) done = !spatial_search(L, aq, xb, yb, db);
    const long i = (long)y * ann_w + x;
    out.xy[i] = PT_XY_TO_INT(xb, yb);
    out.d[i] = db;
    out.done[i] = done;
    out.chg[i] = 1;
}

// same_jump: the previous round used the same jump. A candidate from neighbour n then repeats
// that round's when neither this patch nor n changed in it; it was rejected against the same
// best, so it is skipped (results unchanged).
__global__ __launch_bounds__(THREADS) void k_propagate(LookupArgs L, int J, int same_jump, NnfField cur, NnfField nxt) {
    __shared__ float4 tile[AX * AY];
    stage_tile(L.a, L.ah, L.aw, tile);
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    const int lx = threadIdx.x % TILE_X, ly = threadIdx.x / TILE_X;
    const int x = blockIdx.x * TILE_X + lx, y = blockIdx.y * TILE_Y + ly;
    if (x >= ann_w || y >= ann_h) return;
    const float4 *aq = tile + ly * AX + lx;
    const int bpitch = L.bw * 3, bew = L.bw - 7, beh = L.bh - 7;
    const long i = (long)y * ann_w + x;
    const int v0 = cur.xy[i];
    int xb = PT_INT_TO_X(v0), yb = PT_INT_TO_Y(v0);
    float db = cur.d[i];
    bool moved = false;
    const bool stable = same_jump && !cur.chg[i];
    for (int t = 0; t < 8; t++) {
        int qx = x + J * c_dx[t], qy = y + J * c_dy[t];
        if (qx < 0 || qy < 0 || qx >= ann_w || qy >= ann_h) continue;
        const long qi = (long)qy * ann_w + qx;
        if (stable && !cur.chg[qi]) continue;
        int vq = cur.xy[qi];
        int xs = PT_INT_TO_X(vq) - J * c_dx[t], ys = PT_INT_TO_Y(vq) - J * c_dy[t];
        if (xs < 0 || ys < 0 || xs >= bew || ys >= beh) continue;
        if (xs == xb && ys == yb) continue;
        float d = ssd_staged<true>(aq, L.b + ((size_t)ys * L.bw + xs) * 3, bpitch, db);
        if (d < db) {
            db = d;
            xb = xs;
            yb = ys;
            moved = true;
        }
    }
    bool done = cur.done[i] && !moved;
    bool schg = false;
    if (L.spatial && !done) {
        schg = spatial_search(L, aq, xb, yb, db);
        done = !schg;
    }
    nxt.xy[i] = PT_XY_TO_INT(xb, yb);
    nxt.d[i] = db;
    nxt.done[i] = done;
    nxt.chg[i] = moved || schg;
}

__global__ void k_write_ann(NnfField f, long n, double *__restrict__ ann) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = f.xy[i];
    ann[i * 3 + 0] = (double)PT_INT_TO_X(v);
    ann[i * 3 + 1] = (double)PT_INT_TO_Y(v);
    ann[i * 3 + 2] = (double)f.d[i];
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

int pt_lookup_launch(const LookupArgs &L, NnfField f0, NnfField f1, double *ann_out, hipStream_t s,
                     hipEvent_t ev_after_candidates, hipEvent_t ev_after_rounds) {
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    const long n = (long)ann_w * ann_h;
    dim3 grid((ann_w + TILE_X - 1) / TILE_X, (ann_h + TILE_Y - 1) / TILE_Y);
    hipLaunchKernelGGL(k_candidates, grid, dim3(THREADS), 0, s, L, f0);
    if (ev_after_candidates) PT_HIP_TRY(hipEventRecord(ev_after_candidates, s));
    std::vector<int> J;
    pt_jump_schedule(L.rounds, J);
    for (size_t r = 0; r < J.size(); r++) {
        int same = r > 0 && J[r - 1] == J[r];
        hipLaunchKernelGGL(k_propagate, grid, dim3(THREADS), 0, s, L, J[r], same, f0, f1);
        std::swap(f0, f1);
    }
    if (ev_after_rounds) PT_HIP_TRY(hipEventRecord(ev_after_rounds, s));
    hipLaunchKernelGGL(k_write_ann, dim3(pt_blocks(n, 256)), dim3(256), 0, s, f0, n, ann_out);
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
