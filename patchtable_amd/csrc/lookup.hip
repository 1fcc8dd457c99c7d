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

// ---- Block work queue -------------------------------------------------------------------------
// A lane owns one query patch. Instead of each lane looping over its own candidates (a wave then
// waits for its slowest lane), the lanes of a block push their candidate SSDs to a queue in LDS,
// all lanes of the block evaluate the queue, and each lane reduces its own results in candidate
// order. Every task is bounded by its patch's best distance at the start of the phase, so a
// result is exact whenever it is below that bound; reducing in order with strict "<" then picks
// the same first minimum as evaluating the candidates one after another against the running best.
constexpr int SLOTS = 8;   // tasks per lane and phase

struct BlockQueue {
    unsigned short id[THREADS * SLOTS];   // queued task ids: lane * SLOTS + slot
    int cxy[THREADS * SLOTS];             // candidate XY_TO_INT, by task id
    float res[THREADS * SLOTS];           // SSD, by task id
    float bound[THREADS];                 // per lane
    int n;
};

// Queues the calling lane's task `slot` when `pred`. Called by every lane of the wave (the
// ballot needs them all); one LDS atomic per wave.
__device__ __forceinline__ void q_push(BlockQueue &Q, bool pred, int slot, int cand) {
    const unsigned long long m = __ballot(pred);
    if (m == 0) return;
    const int wl = threadIdx.x & 63;
    const int leader = __builtin_ctzll(m);
    int base = 0;
    if (wl == leader) base = atomicAdd(&Q.n, __popcll(m));
    base = __shfl(base, leader);
    if (pred) {
        const int pos = base + __popcll(m & ((1ull << wl) - 1));
        Q.id[pos] = (unsigned short)(threadIdx.x * SLOTS + slot);
        Q.cxy[threadIdx.x * SLOTS + slot] = cand;
    }
}

// Evaluates every queued task; on return the results are visible to all lanes and the queue is
// empty again.
__device__ __forceinline__ void q_run(BlockQueue &Q, const float4 *tile, const LookupArgs &L) {
    __syncthreads();
    const int n = Q.n;
    const int bpitch = L.bw * 3;
    for (int k = threadIdx.x; k < n; k += THREADS) {
        const int id = Q.id[k];
        const int lane = id / SLOTS;
        const int c = Q.cxy[id];
        const float4 *aq = tile + (lane / TILE_X) * AX + (lane % TILE_X);
        Q.res[id] = ssd_staged<true>(aq, L.b + ((size_t)PT_INT_TO_Y(c) * L.bw + PT_INT_TO_X(c)) * 3, bpitch,
                                     Q.bound[lane]);
    }
    __syncthreads();
    if (threadIdx.x == 0) Q.n = 0;
    __syncthreads();
}

// Spatial search (patchtable.h:3861-3885) around (xb, yb), including the reference's "else if"
// clamping; the far side is also clamped where the reference would read out of bounds (a
// database fewer than 2r+1 patches tall or wide). Window positions are visited row by row.
struct Window {
    int ymin, ymax, xmin, xmax, xb0, yb0;
};

__device__ __forceinline__ Window spatial_window(const LookupArgs &L, int xb, int yb) {
    const int bew = L.bw - 7, beh = L.bh - 7, r = L.spatial;
    Window w{yb - r, yb + r + 1, xb - r, xb + r + 1, xb, yb};
    if (w.ymin < 0) w.ymin = 0;
    else if (w.ymax > beh) w.ymax = beh;
    if (w.xmin < 0) w.xmin = 0;
    else if (w.xmax > bew) w.xmax = bew;
    if (w.ymax > beh) w.ymax = beh;
    if (w.xmax > bew) w.xmax = bew;
    return w;
}

// Sequential form, for spatial radius > 1 (more than SLOTS positions). Returns true if improved.
__device__ __forceinline__ bool spatial_search_seq(const LookupArgs &L, const float4 *aq, int &xb, int &yb, float &db) {
    const Window w = spatial_window(L, xb, yb);
    const int bpitch = L.bw * 3;
    bool changed = false;
    for (int ys = w.ymin; ys < w.ymax; ys++)
        for (int xs = w.xmin; xs < w.xmax; xs++) {
            if (xs == w.xb0 && ys == w.yb0) continue;
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

// Spatial search through the block queue (radius 1: at most 8 positions). `want` is false for
// lanes that skip it; every lane of the block must call this. Returns true if improved.
__device__ __forceinline__ bool spatial_search_q(BlockQueue &Q, const float4 *tile, const LookupArgs &L, bool want,
                                                 int &xb, int &yb, float &db) {
    Window w = spatial_window(L, xb, yb);
    if (!want) w.ymax = w.ymin;   // empty
    if (want) Q.bound[threadIdx.x] = db;
    int slot = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            if (dx == 0 && dy == 0) continue;
            // slot order = row-by-row order of the window positions other than the centre
            const int ys = w.yb0 + dy, xs = w.xb0 + dx;
            const bool in = ys >= w.ymin && ys < w.ymax && xs >= w.xmin && xs < w.xmax;
            q_push(Q, in, slot, PT_XY_TO_INT(xs, ys));
            slot++;
        }
    q_run(Q, tile, L);
    bool changed = false;
    if (want) {
        int s = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                if (dx == 0 && dy == 0) continue;
                const int ys = w.yb0 + dy, xs = w.xb0 + dx;
                if (ys >= w.ymin && ys < w.ymax && xs >= w.xmin && xs < w.xmax) {
                    const float d = Q.res[threadIdx.x * SLOTS + s];
                    if (d < db) {
                        db = d;
                        xb = xs;
                        yb = ys;
                        changed = true;
                    }
                }
                s++;
            }
    }
    return changed;
}

__global__ __launch_bounds__(THREADS) void k_candidates(LookupArgs L, NnfField out) {
    __shared__ float4 tile[AX * AY];
    __shared__ BlockQueue Q;
    if (threadIdx.x == 0) Q.n = 0;
    stage_tile(L.a, L.ah, L.aw, tile);
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    const int lx = threadIdx.x % TILE_X, ly = threadIdx.x / TILE_X;
    const int x = blockIdx.x * TILE_X + lx, y = blockIdx.y * TILE_Y + ly;
    const bool valid = x < ann_w && y < ann_h;
    const float4 *aq = tile + ly * AX + lx;
    const int bpitch = L.bw * 3, bew = L.bw - 7;
    int xb = 0, yb = 0;
    float db = 0.f;
    if (valid) {
        // table cell of the query descriptor at (y+7, x+7)
        const size_t plane = (size_t)(L.ah + 7) * (L.aw + 7);
        const size_t off = (size_t)(y + 7) * (L.aw + 7) + x + 7;
        int64_t cell = 0;
        for (int d = 0; d < L.q.G; d++)
            cell += (int64_t)pt_get_bin(L.a_desc[d * plane + off], L.q.min_val[d], L.q.range[d],
                                        L.q.to_bin + (size_t)d * PT_NMAP) *
                    L.q.stride[d];
        const int v = L.table[cell];
        xb = PT_INT_TO_X(v) - 7;
        yb = PT_INT_TO_Y(v) - 7;
        db = ssd_staged<false>(aq, L.b + ((size_t)yb * L.bw + xb) * 3, bpitch, 0.f);
    }
    if (L.K > 0 && L.kcoh_iter0) {
        // The k-coherence list of the table candidate is walked in order and abandoned at the
        // first entry with kd >= triangle_factor * dbest (patchtable.h:3779-3784); kd is sorted,
        // so per chunk of SLOTS entries the queued superset is those with kd below the bound at
        // the chunk's start, and the in-order reduce replays the early-out exactly.
        const size_t lst = ((size_t)yb * bew + xb) * L.K;
        bool broken = !valid;
        for (int k0 = 0; k0 < L.K; k0 += SLOTS) {
            const double tb = L.triangle_factor * (double)db;
            if (!broken) Q.bound[threadIdx.x] = db;
            for (int s = 0; s < SLOTS; s++) {
                const int k = k0 + s;
                bool pred = !broken && k < L.K && (double)L.kdist[lst + k] < tb;
                int pos = pred ? L.kpos[lst + k] : 0;
                q_push(Q, pred, s, pos);
            }
            q_run(Q, tile, L);
            for (int s = 0; s < SLOTS && !broken; s++) {
                const int k = k0 + s;
                if (k >= L.K) break;
                if ((double)L.kdist[lst + k] >= L.triangle_factor * (double)db) {
                    broken = true;
                    break;
                }
                const float d = Q.res[threadIdx.x * SLOTS + s];
                if (d < db) {
                    db = d;
                    const int c = Q.cxy[threadIdx.x * SLOTS + s];
                    xb = PT_INT_TO_X(c);
                    yb = PT_INT_TO_Y(c);
                }
            }
        }
    }
    bool done = false;
    if (L.spatial == 1) done = !spatial_search_q(Q, tile, L, valid, xb, yb, db);
    else if (L.spatial > 1 && valid) done = !spatial_search_seq(L, aq, xb, yb, db);
    if (!valid) return;
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
    __shared__ BlockQueue Q;
    if (threadIdx.x == 0) Q.n = 0;
    stage_tile(L.a, L.ah, L.aw, tile);
    const int ann_w = L.aw - 7, ann_h = L.ah - 7;
    const int lx = threadIdx.x % TILE_X, ly = threadIdx.x / TILE_X;
    const int x = blockIdx.x * TILE_X + lx, y = blockIdx.y * TILE_Y + ly;
    const bool valid = x < ann_w && y < ann_h;
    const float4 *aq = tile + ly * AX + lx;
    const int bew = L.bw - 7, beh = L.bh - 7;
    const long i = (long)y * ann_w + x;
    int xb = 0, yb = 0;
    float db = 0.f;
    bool stable = false, was_done = false;
    if (valid) {
        const int v0 = cur.xy[i];
        xb = PT_INT_TO_X(v0);
        yb = PT_INT_TO_Y(v0);
        db = cur.d[i];
        stable = same_jump && !cur.chg[i];
        was_done = cur.done[i];
        Q.bound[threadIdx.x] = db;
    }
    unsigned pushed = 0;
    for (int t = 0; t < 8; t++) {
        const int qx = x + J * c_dx[t], qy = y + J * c_dy[t];
        bool pred = valid && qx >= 0 && qy >= 0 && qx < ann_w && qy < ann_h;
        int xs = 0, ys = 0;
        if (pred) {
            const long qi = (long)qy * ann_w + qx;
            if (stable && !cur.chg[qi]) pred = false;
            else {
                const int vq = cur.xy[qi];
                xs = PT_INT_TO_X(vq) - J * c_dx[t];
                ys = PT_INT_TO_Y(vq) - J * c_dy[t];
                pred = xs >= 0 && ys >= 0 && xs < bew && ys < beh && !(xs == xb && ys == yb);
            }
        }
        if (pred) pushed |= 1u << t;
        q_push(Q, pred, t, PT_XY_TO_INT(xs, ys));
    }
    q_run(Q, tile, L);
    bool moved = false;
    for (int t = 0; t < 8; t++) {
        if (!((pushed >> t) & 1u)) continue;
        const float d = Q.res[threadIdx.x * SLOTS + t];
        if (d < db) {
            db = d;
            const int c = Q.cxy[threadIdx.x * SLOTS + t];
            xb = PT_INT_TO_X(c);
            yb = PT_INT_TO_Y(c);
            moved = true;
        }
    }
    bool done = was_done && !moved;
    bool schg = false;
    if (L.spatial == 1) {
        schg = spatial_search_q(Q, tile, L, valid && !done, xb, yb, db);
        if (valid && !done) done = !schg;
    } else if (L.spatial > 1 && valid && !done) {
        schg = spatial_search_seq(L, aq, xb, yb, db);
        done = !schg;
    }
    if (!valid) return;
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
