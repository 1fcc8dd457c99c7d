// Query lookup on gfx950 (replaces lookup_templated<is_descriptor, use_allowed_patches>,
// patchtable.h:3468-4217).
//
// The reference visits the query patches of the lattice {x, y multiples of query_step} in raster
// order (then reverse raster order) and lets each visited patch push its best offset to its
// (2*query_step+1)^2 neighbourhood before those are visited, so a good match propagates along a
// whole row within one pass. Here every lattice patch is processed in parallel and propagation
// runs in rounds over double-buffered fields:
//   init: nothing, or the previous field (clamped; distance recomputed and scaled by the temporal
//     weight, or the previous distance scaled, patchtable.h:3539-3585);
//   candidates: the cell of every table (patchtable.h:3664-3737), the k-coherence list of the
//     best with the triangle early-out (patchtable.h:3748-3851), and one random neighbour (do_rs
//     without spatial, patchtable.h:3886-3898);
//   rounds: jump flooding -- the 8 lattice neighbours at distance J (J = 8, 4, 2, 1, then 1 ...)
//     each propose their offset; the k-coherence walk again when kcoherence_iter asks for it after
//     the first pass or when there are no table candidates; then, from the second round on, the
//     spatial window around the best (patchtable.h:3853-3885, allowed positions only), or a random
//     neighbour. (The window runs in the candidate pass only when there are no rounds, and in the
//     first round only when there is one round: searching it around the early, still-moving
//     matches costs more than it gains);
//   gather (query_step > 1): every patch takes the best of the offsets of the lattice patches
//     within query_step of it, as the reference's propagation scatter does;
//   output: NNF_X, NNF_Y, NNF_DIST as doubles, with the temporal rescaling of patchtable.h:4178-4193.
// Candidates replace the best only when strictly closer and every distance is exact (the RGB SSD
// with the reference's lane order, or patch_dist_approx of the descriptors in descriptor mode),
// so NNF_DIST is always the distance of the stored match.
//
// Work layout: a workgroup of 256 lanes owns a 64 x 4 tile of lattice patches. With
// query_step 1 in image mode the tile's 71 x 11 query pixels are staged in LDS (one float4 per
// pixel); in descriptor mode each lane's query descriptor is staged; with query_step > 1 the
// query patches are read from global memory. The field is packed as XY_TO_INT position +
// distance + flags: the spatial window around the stored best was already searched without
// improvement (a later round whose best is unchanged skips it), the best changed in the round
// that produced the field (used to skip repeated candidates), and the k-coherence list of the
// best was already walked. None of these skips changes a result.
#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_tile.h"

#include <cmath>
#include <type_traits>
#include <vector>

namespace {

using namespace pt_tile;

__constant__ int c_dx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int c_dy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

// Where a lane's query patch is read from: the float tile in LDS, global memory, the descriptor
// in LDS; M_INT8 (MFMA form's candidate pass): the int8 RGB0 tile in LDS, distances being int8
// SSDs |q|^2 + |p|^2 - 2 q.p (exact integers) instead of FP32 SSDs.
enum { M_TILE = 0, M_GLOBAL = 1, M_DESC = 2, M_INT8 = 3 };

constexpr uint8_t F_DONE = 1, F_CHG = 2, F_KD = 4, F_MOVED = PT_FIELD_MOVED;

// Floats of LDS that hold the block's query patches.
template <int M, int TX>
struct QBuf {
    static constexpr int F4 = M == M_TILE   ? (TX + 7) * (THREADS / TX + 7)
                              : M == M_INT8 ? ((TX + 7) * (THREADS / TX + 7) + 3) / 4
                              : M == M_DESC ? THREADS * PT_LOOKUP_CMAX / 4
                                            : 1;
};

template <int TX>
__device__ __forceinline__ void lane_xy(const LookupArgs &L, int lane, int &x, int &y) {
    const int2 tc = tile_coords();
    x = (tc.x * TX + lane % TX) * L.qs;
    y = (tc.y * (THREADS / TX) + lane / TX) * L.qs;
}

// Stages the block's query patches (every lane of the block must call this).
template <int M, int TX>
__device__ __forceinline__ void stage_query(const LookupArgs &L, float4 *qbuf, bool valid, int x, int y) {
    if constexpr (M == M_TILE) {
        stage_tile<TX>(L.a, L.ah, L.aw, qbuf);
    } else if constexpr (M == M_INT8) {
        constexpr int SX = TX + 7, SY = THREADS / TX + 7;
        uint32_t *t = (uint32_t *)qbuf;
        const int2 tc = tile_coords();
        const int x0 = tc.x * TX, y0 = tc.y * (THREADS / TX);
        for (int i = threadIdx.x; i < SX * SY; i += THREADS) {
            const int ly = i / SX, lx = i - ly * SX, gx = x0 + lx, gy = y0 + ly;
            t[i] = gx < L.aw && gy < L.ah ? L.aq8[(size_t)gy * L.aw + gx] : 0u;
        }
        __syncthreads();
    } else if constexpr (M == M_DESC) {
        float *qd = (float *)qbuf + threadIdx.x * L.C;
        const float *src = L.a_descI + ((size_t)(y + 7) * (L.aw + 7) + x + 7) * L.C;
        for (int c = 0; c < L.C; c++) qd[c] = valid ? src[c] : 0.f;
        __syncthreads();
    }
}

// Distance of lane `lane`'s query patch to database patch (xs, ys); with a finite bound the SSD
// may stop early at a partial sum >= bound (then the result is >= bound, never a smaller value).
template <int M, int TX>
__device__ __forceinline__ float qdist(const LookupArgs &L, const float4 *qbuf, int lane, int xs, int ys, float bound) {
    if constexpr (M == M_TILE) {
        const float4 *aq = qbuf + (lane / TX) * (TX + 7) + (lane % TX);
        return ssd_staged<true, TX + 7>(aq, L.b + ((size_t)ys * L.bw + xs) * 3, L.bw * 3, bound);
    } else if constexpr (M == M_INT8) {
        // |q|^2 + |p|^2 - 2 q.p over the RGB0 words (the fourth byte is zero in both)
        constexpr int SX = TX + 7;
        const uint32_t *qt = (const uint32_t *)qbuf + (lane / TX) * SX + (lane % TX);
        const int pitch = L.bw + 2 * PT_MFMA_PAD;
        const int32_t *bp = (const int32_t *)L.bq8 + (size_t)(ys + PT_MFMA_PAD) * pitch + xs + PT_MFMA_PAD;
        int dot = 0;
#pragma unroll 2
        for (int r = 0; r < 8; r++) {
            const int4 b0 = ldg4i(bp + (size_t)r * pitch), b1 = ldg4i(bp + (size_t)r * pitch + 4);
            const uint32_t *q = qt + r * SX;
            dot = __builtin_amdgcn_sdot4((int)q[0], b0.x, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[1], b0.y, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[2], b0.z, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[3], b0.w, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[4], b1.x, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[5], b1.y, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[6], b1.z, dot, false);
            dot = __builtin_amdgcn_sdot4((int)q[7], b1.w, dot, false);
        }
        int x, y;
        lane_xy<TX>(L, lane, x, y);
        const int nkp = L.bw - 7 + 2 * PT_MFMA_PAD;
        const int pn = (int)(L.bnk[(size_t)(ys + PT_MFMA_PAD) * nkp + xs + PT_MFMA_PAD] >> 7) - (int)PT_MFMA_OFF;
        return (float)(L.qnorm[(size_t)y * L.ann_w + x] + pn - 2 * dot);
    } else if constexpr (M == M_GLOBAL) {
        int x, y;
        lane_xy<TX>(L, lane, x, y);
        return pt_patch_ssd_bounded(L.a + ((size_t)y * L.aw + x) * 3, L.aw * 3, L.b + ((size_t)ys * L.bw + xs) * 3,
                                    L.bw * 3, bound);
    } else {
        const float *qd = (const float *)qbuf + lane * L.C;
        return pt_desc_dist(qd, L.b_descI + ((size_t)(ys + 7) * (L.bw + 7) + xs + 7) * L.C, L.C);
    }
}

__device__ __forceinline__ bool allowed_at(const LookupArgs &L, int xs, int ys) {
    return !L.allowed || L.allowed[(size_t)ys * (L.bw - 7) + xs] == L.allowed_index;
}

// Candidates already evaluated for this patch, in this lookup: the last HIST of them, in a ring.
// Evaluating one again cannot change the result (it was rejected against a best distance that
// has only decreased since, or it became the best and was replaced by a closer one, or it is the
// best), so such repeats are skipped; keeping any subset of the evaluated candidates is exact.
// Between the kernels of a lookup the ring lives in HBM, entry-major (entry k of lattice patch v
// at hist[k * nv + v]) with the count of recorded candidates in hn[v].
constexpr int HIST = PT_LOOKUP_HIST;
struct Hist {
    int c[HIST];
    int n = 0;
    int adds = 0;   // candidates recorded since the last load
    __device__ __forceinline__ bool has(int v) const {
        bool f = false;
#pragma unroll
        for (int i = 0; i < HIST; i++) f |= i < n && c[i] == v;
        return f;
    }
    __device__ __forceinline__ void add(int v) {
        const int k = n % HIST;
#pragma unroll
        for (int i = 0; i < HIST; i++)
            if (i == k) c[i] = v;
        n = n + 1 >= 2 * HIST ? n + 1 - HIST : n + 1;   // n % HIST is the next slot; n >= HIST once full
        adds++;
    }
    __device__ __forceinline__ void load(const LookupArgs &L, long v, long nv) {
        n = L.hn[v];
        adds = 0;
#pragma unroll
        for (int i = 0; i < HIST; i++) c[i] = i < n ? L.hist[i * nv + v] : 0;
    }
    // Writes the count and the entries recorded since the load (the newest `adds` slots).
    __device__ __forceinline__ void store(const LookupArgs &L, long v, long nv) const {
        if (adds == 0) return;
        L.hn[v] = (uint8_t)n;
        const int added = adds < HIST ? adds : HIST;
#pragma unroll
        for (int i = 0; i < HIST; i++) {
            const int age = ((n - 1 - i) % HIST + HIST) % HIST;   // 0: the newest slot
            if (i < n && age < added) L.hist[i * nv + v] = c[i];
        }
    }
};

// ---- Block work queue -------------------------------------------------------------------------
// A lane owns one query patch. Instead of each lane looping over its own candidates (a wave then
// waits for its slowest lane), the lanes of a block push their candidate distances to a queue in
// LDS, all lanes of the block evaluate the queue, and each lane reduces its own results in
// candidate order. Every task is bounded by its patch's best distance at the start of the phase,
// so a result is exact whenever it is below that bound; reducing in order with strict "<" then
// picks the same first minimum as evaluating the candidates one after another.
constexpr int SLOTS = 8;   // tasks per lane and phase

struct BlockQueue {
    unsigned short id[THREADS * SLOTS];   // queued task ids: lane * SLOTS + slot
    int cxy[THREADS * SLOTS];             // candidate XY_TO_INT, by task id
    float res[THREADS * SLOTS];           // distance, by task id
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
template <int M, int TX>
__device__ __forceinline__ void q_run(BlockQueue &Q, const float4 *qbuf, const LookupArgs &L) {
    __syncthreads();
    const int n = Q.n;
    for (int k = threadIdx.x; k < n; k += THREADS) {
        const int id = Q.id[k];
        const int lane = id / SLOTS;
        const int c = Q.cxy[id];
        Q.res[id] = qdist<M, TX>(L, qbuf, lane, PT_INT_TO_X(c), PT_INT_TO_Y(c), Q.bound[lane]);
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
template <int M, int TX>
__device__ __forceinline__ bool spatial_search_seq(const LookupArgs &L, const float4 *qbuf, int &xb, int &yb, float &db,
                                                   Hist &hist) {
    const Window w = spatial_window(L, xb, yb);
    bool changed = false;
    for (int ys = w.ymin; ys < w.ymax; ys++)
        for (int xs = w.xmin; xs < w.xmax; xs++) {
            if ((xs == w.xb0 && ys == w.yb0) || !allowed_at(L, xs, ys) || hist.has(PT_XY_TO_INT(xs, ys))) continue;
            hist.add(PT_XY_TO_INT(xs, ys));
            float d = qdist<M, TX>(L, qbuf, threadIdx.x, xs, ys, db);
            if (d < db) {
                db = d;
                xb = xs;
                yb = ys;
                changed = true;
            }
        }
    return changed;
}

// The 3 x 3 window around an interior best (1 <= xb <= bew-2, 1 <= yb <= beh-2) evaluated by
// the calling lane alone, for the candidates in `todo` (slot order: row by row, centre left
// out). The candidates of one window row (dy) read their 8 patch rows together: 30 floats, 8
// loads per row instead of 6 per candidate row. Each candidate keeps the four lane-order
// accumulators of ssd_staged in its order and stops once a partial sum after a row other than
// the last reaches `bound`, so res[slot] is what qdist returns for it.
template <int TX>
__device__ __forceinline__ void spatial_local(const LookupArgs &L, const float4 *qbuf, int xb, int yb, float bound,
                                              unsigned todo, float res[8]) {
    const float4 *aq = qbuf + (threadIdx.x / TX) * (TX + 7) + (threadIdx.x % TX);
#pragma unroll
    for (int g = 0; g < 3; g++) {
        const int s0 = g == 0 ? 0 : (g == 1 ? 3 : 5);   // first slot of window row dy = g - 1
        const int nc = g == 1 ? 2 : 3;                  // the centre row holds dx = -1 and +1
        unsigned alive = (todo >> s0) & ((1u << nc) - 1u);
        if (!alive) continue;
        float acc[3][4];
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) acc[j][e] = 0.f;
        const float *bw0 = L.b + ((size_t)(yb + g - 1) * L.bw + (xb - 1)) * 3;
#pragma unroll 1
        for (int r = 0; r < 8 && alive; r++) {
            const float *br = bw0 + (size_t)r * L.bw * 3;
            float wv[30];
#pragma unroll
            for (int k = 0; k < 7; k++) {
                const float4 v = ldg4(br + 4 * k);
                wv[4 * k] = v.x;
                wv[4 * k + 1] = v.y;
                wv[4 * k + 2] = v.z;
                wv[4 * k + 3] = v.w;
            }
            wv[28] = br[28];
            wv[29] = br[29];
            float av[24];
            const float4 *ar = aq + r * (TX + 7);
#pragma unroll
            for (int px = 0; px < 8; px++) {
                const float4 p = ar[px];
                av[3 * px] = p.x;
                av[3 * px + 1] = p.y;
                av[3 * px + 2] = p.z;
            }
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (j >= nc || !((alive >> j) & 1u)) continue;
                const int off = g == 1 ? 6 * j : 3 * j;   // 3 * (dx + 1)
#pragma unroll
                for (int e = 0; e < 24; e++) {
                    const float d = av[e] - wv[off + e];
                    acc[j][e & 3] = acc[j][e & 3] + d * d;
                }
                const float part = ((acc[j][0] + acc[j][1]) + acc[j][2]) + acc[j][3];
                if (r == 7 || part >= bound) {
                    res[s0 + j] = part;
                    alive &= ~(1u << j);
                }
            }
        }
    }
}

// Spatial search (radius 1: at most 8 positions). `want` is false for lanes that skip it; every
// lane of the block must call this. Returns true if improved. In image mode with query_step 1,
// when most lanes of the wave search, a lane whose window lies inside b evaluates it alone
// (spatial_local); the others go through the block queue.
template <int M, int TX, bool LOCAL>
__device__ __forceinline__ bool spatial_search_q(BlockQueue &Q, const float4 *qbuf, const LookupArgs &L, bool want,
                                                 int &xb, int &yb, float &db, Hist &hist) {
    Window w = spatial_window(L, xb, yb);
    if (!want) w.ymax = w.ymin;   // empty
    if (want) Q.bound[threadIdx.x] = db;
    bool local = false;
    if constexpr (M == M_TILE && LOCAL) {
        const bool interior = want && xb >= 1 && yb >= 1 && xb <= L.bw - 9 && yb <= L.bh - 9;
        local = interior && __popcll(__ballot(want)) >= 32;
    }
    unsigned pushed = 0;
    int slot = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            if (dx == 0 && dy == 0) continue;
            // slot order = row-by-row order of the window positions other than the centre
            const int ys = w.yb0 + dy, xs = w.xb0 + dx;
            const bool in = ys >= w.ymin && ys < w.ymax && xs >= w.xmin && xs < w.xmax && allowed_at(L, xs, ys) &&
                            !hist.has(PT_XY_TO_INT(xs, ys));
            if (in) {
                pushed |= 1u << slot;
                hist.add(PT_XY_TO_INT(xs, ys));
            }
            q_push(Q, in && !local, slot, PT_XY_TO_INT(xs, ys));
            slot++;
        }
    q_run<M, TX>(Q, qbuf, L);
    float lres[8];
    if constexpr (M == M_TILE && LOCAL)
        if (local && pushed) spatial_local<TX>(L, qbuf, xb, yb, db, pushed, lres);
    bool changed = false;
    for (int s = 0; s < 8; s++) {
        if (!((pushed >> s) & 1u)) continue;
        const float d = local ? lres[s] : Q.res[threadIdx.x * SLOTS + s];
        if (d < db) {
            db = d;
            // slot s sits at (c_dx[s], c_dy[s]) from the window's centre
            const int c = local ? PT_XY_TO_INT(w.xb0 + c_dx[s], w.yb0 + c_dy[s]) : Q.cxy[threadIdx.x * SLOTS + s];
            xb = PT_INT_TO_X(c);
            yb = PT_INT_TO_Y(c);
            changed = true;
        }
    }
    return changed;
}

// Walk of the k-coherence list of the best at the walk's start, abandoned at the first entry with
// kd >= triangle_factor * dbest (patchtable.h:3779-3784). kd is sorted, so per chunk of SLOTS
// entries the queued superset is those with kd below the bound at the chunk's start, and the
// in-order reduce replays the early-out exactly. Every lane of the block must call this.
template <int M, int TX>
__device__ __forceinline__ bool kcoh_walk_q(BlockQueue &Q, const float4 *qbuf, const LookupArgs &L, bool want, int &xb,
                                            int &yb, float &db, Hist &hist) {
    const int bew = L.bw - 7;
    const size_t lst = ((size_t)yb * bew + xb) * L.K;
    bool broken = !want, changed = false;
    for (int k0 = 0; k0 < L.K; k0 += SLOTS) {
        const double tb = L.triangle_factor * (double)db;
        if (!broken) Q.bound[threadIdx.x] = db;
        unsigned pushed = 0;
        for (int s = 0; s < SLOTS; s++) {
            const int k = k0 + s;
            bool pred = !broken && k < L.K && (double)L.kdist[lst + k] < tb;
            int pos = pred ? L.kpos[lst + k] : 0;
            if (pred && hist.has(pos)) pred = false;   // evaluated before: cannot win
            if (pred) pushed |= 1u << s;
            q_push(Q, pred, s, pos);
        }
        q_run<M, TX>(Q, qbuf, L);
        for (int s = 0; s < SLOTS && !broken; s++) {
            const int k = k0 + s;
            if (k >= L.K) break;
            if ((double)L.kdist[lst + k] >= L.triangle_factor * (double)db) {
                broken = true;
                break;
            }
            if (!((pushed >> s) & 1u)) continue;
            // recorded only once its result is used: an entry past the walk's end was evaluated
            // but never compared, so it may still win later
            hist.add(Q.cxy[threadIdx.x * SLOTS + s]);
            const float d = Q.res[threadIdx.x * SLOTS + s];
            if (d < db) {
                db = d;
                const int c = Q.cxy[threadIdx.x * SLOTS + s];
                xb = PT_INT_TO_X(c);
                yb = PT_INT_TO_Y(c);
                changed = true;
            }
        }
    }
    return changed;
}

// Random neighbour of the random search (do_rs without spatial): a hash of the patch index and
// the round replaces the reference's rand() & 7 (patchtable.h:3887).
__device__ __forceinline__ int rs_pick(uint32_t i, uint32_t round) {
    uint32_t h = i * 0x9E3779B1u + round * 0x85EBCA77u + 0x165667B1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return (int)(h & 7u);
}

template <int M, int TX>
__device__ __forceinline__ bool random_search(const LookupArgs &L, const float4 *qbuf, long pi, int round, int &xb,
                                              int &yb, float &db) {
    const int r = rs_pick((uint32_t)pi, (uint32_t)round);
    const int xs = xb + c_dx[r], ys = yb + c_dy[r];
    if (xs < 0 || ys < 0 || xs >= L.bw - 7 || ys >= L.bh - 7 || !allowed_at(L, xs, ys)) return false;
    const float d = qdist<M, TX>(L, qbuf, threadIdx.x, xs, ys, db);
    if (d < db) {
        db = d;
        xb = xs;
        yb = ys;
        return true;
    }
    return false;
}

__device__ __forceinline__ int table_cell_value(const LookupArgs &L, const float *desc, const DevQuant &q,
                                                const int32_t *table, int x, int y) {
    const size_t plane = (size_t)(L.ah + 7) * (L.aw + 7);
    const size_t off = (size_t)(y + 7) * (L.aw + 7) + x + 7;
    int64_t cell = 0;
    for (int d = 0; d < q.G; d++)
        cell += (int64_t)pt_get_bin(desc[d * plane + off], q.min_val[d], q.range[d], q.to_bin + (size_t)d * PT_NMAP) *
                q.stride[d];
    return table[cell];
}

// The initial match (a previous field's, scored by distance or, in M_INT8, by int8 SSD times the
// temporal weight) and the cells of every table (patchtable.h:3664-3737): a cell equal to the
// current match, or to the previous table's cell, is skipped; the others replace the match when
// strictly closer. `moved`: the match is not the initial one.
template <int M, int TX, bool REC = true>
__device__ __forceinline__ void init_and_tables(const LookupArgs &L, const float4 *qbuf, bool valid, int x, int y,
                                                long pi, bool &have, bool &moved, int &xb, int &yb, float &db,
                                                Hist &hist) {
    have = false;
    moved = true;
    xb = 0;
    yb = 0;
    db = 0.f;
    if (valid && L.init_xy) {
        const int v = L.init_xy[pi];
        xb = PT_INT_TO_X(v);
        yb = PT_INT_TO_Y(v);
        if constexpr (M == M_INT8)
            db = (float)((double)qdist<M, TX>(L, qbuf, threadIdx.x, xb, yb, INFINITY) *
                         (L.init_sv ? (double)L.init_sv[pi] : L.init_ts));
        else
            db = L.init_d[pi];
        have = true;
        moved = false;
    }
    if (valid && L.do_table_lookup) {
        const int v = table_cell_value(L, L.a_desc, L.q, L.table, x, y);
        const int xs = PT_INT_TO_X(v) - 7, ys = PT_INT_TO_Y(v) - 7;
        if (!have || xs != xb || ys != yb) {
            if (REC) hist.add(PT_XY_TO_INT(xs, ys));
            const float d = qdist<M, TX>(L, qbuf, threadIdx.x, xs, ys, have ? db : INFINITY);
            if (!have || d < db) {
                db = d;
                xb = xs;
                yb = ys;
                moved = true;
            }
            have = true;
        }
        int xp = xs, yp = ys;
        for (int t = 0; t < L.nxt; t++) {   // further tables (patchtable.h:3711-3737)
            const int v1 = table_cell_value(L, L.xa_desc, L.xq[t], L.xtable[t], x, y);
            const int x1 = PT_INT_TO_X(v1) - 7, y1 = PT_INT_TO_Y(v1) - 7;
            if ((x1 != xb || y1 != yb) && (x1 != xp || y1 != yp)) {
                if (REC) hist.add(PT_XY_TO_INT(x1, y1));
                const float d = qdist<M, TX>(L, qbuf, threadIdx.x, x1, y1, db);
                if (d < db) {
                    db = d;
                    xb = x1;
                    yb = y1;
                    moved = true;
                }
            }
            xp = x1;
            yp = y1;
        }
    }
}

template <int M, int TX>
__device__ __forceinline__ void candidates_body(const LookupArgs &L, NnfField out) {
    __shared__ float4 qbuf[QBuf<M, TX>::F4];
    __shared__ BlockQueue Q;
    if (threadIdx.x == 0) Q.n = 0;
    const int2 tc = tile_coords();
    const int vx = tc.x * TX + threadIdx.x % TX, vy = tc.y * (THREADS / TX) + threadIdx.x / TX;
    const bool valid = vx < L.vw && vy < L.vh;
    const int x = vx * L.qs, y = vy * L.qs;
    stage_query<M, TX>(L, qbuf, valid, x, y);
    const long pi = (long)y * L.ann_w + x;
    bool have, moved;
    int xb, yb;
    float db;
    Hist hist;
    init_and_tables<M, TX>(L, qbuf, valid, x, y, pi, have, moved, xb, yb, db, hist);
    bool kd = false;
    if (L.K > 0 && L.kcoh_iter0) kd = !kcoh_walk_q<M, TX>(Q, qbuf, L, valid, xb, yb, db, hist);
    bool done = false;
    if (L.spatial_cand && L.spatial == 1) {
        const bool changed = spatial_search_q<M, TX, false>(Q, qbuf, L, valid, xb, yb, db, hist);
        done = !changed;
        if (changed) kd = false;
    } else if (L.spatial_cand && L.spatial > 1 && valid) {
        const bool changed = spatial_search_seq<M, TX>(L, qbuf, xb, yb, db, hist);
        done = !changed;
        if (changed) kd = false;
    } else if (L.rs && valid) {
        if (random_search<M, TX>(L, qbuf, pi, 0, xb, yb, db)) kd = false;
    }
    if (!valid) return;
    const long v = (long)vy * L.vw + vx;
    out.xy[v] = PT_XY_TO_INT(xb, yb);
    out.d[v] = db;
    out.fl[v] = (done ? F_DONE : 0) | F_CHG | (kd ? F_KD : 0);
    if (L.hist && L.rounds > 0) {
        if (hist.adds == 0) L.hn[v] = 0;   // nothing evaluated: the count must still be written
        hist.store(L, v, (long)L.vw * L.vh);
    }
}

// same_jump: the previous round used the same jump. A candidate from neighbour n then repeats
// that round's when neither this patch nor n changed in it; it was rejected against the same
// best, so it is skipped (results unchanged).
template <int M, int TX>
__device__ __forceinline__ void propagate_body(const LookupArgs &L, int J, int same_jump, int round, NnfField cur,
                                               NnfField nxt, int keep_hist) {
    __shared__ float4 qbuf[QBuf<M, TX>::F4];
    __shared__ BlockQueue Q;
    if (threadIdx.x == 0) Q.n = 0;
    const int2 tc = tile_coords();
    const int vx = tc.x * TX + threadIdx.x % TX, vy = tc.y * (THREADS / TX) + threadIdx.x / TX;
    const bool valid = vx < L.vw && vy < L.vh;
    const int x = vx * L.qs, y = vy * L.qs;
    stage_query<M, TX>(L, qbuf, valid, x, y);
    const int bew = L.bw - 7, beh = L.bh - 7;
    const long v = (long)vy * L.vw + vx;
    const int js = J * L.qs;   // jump in patch units
    int xb = 0, yb = 0;
    float db = 0.f;
    bool stable = false, was_done = false, kd = false;
    const long nv = (long)L.vw * L.vh;
    Hist hist;
    if (valid) {
        if (L.hist) hist.load(L, v, nv);
        const int v0 = cur.xy[v];
        xb = PT_INT_TO_X(v0);
        yb = PT_INT_TO_Y(v0);
        db = cur.d[v];
        const uint8_t fl = cur.fl[v];
        stable = same_jump && !(fl & F_CHG);
        was_done = fl & F_DONE;
        kd = fl & F_KD;
        Q.bound[threadIdx.x] = db;
    }
    unsigned pushed = 0;
    for (int t = 0; t < 8; t++) {
        const int qx = vx + J * c_dx[t], qy = vy + J * c_dy[t];
        bool pred = valid && qx >= 0 && qy >= 0 && qx < L.vw && qy < L.vh;
        int xs = 0, ys = 0;
        if (pred) {
            const long qv = (long)qy * L.vw + qx;
            if (stable && !(cur.fl[qv] & F_CHG)) pred = false;
            else {
                const int vq = cur.xy[qv];
                xs = PT_INT_TO_X(vq) - js * c_dx[t];
                ys = PT_INT_TO_Y(vq) - js * c_dy[t];
                pred = xs >= 0 && ys >= 0 && xs < bew && ys < beh && !(xs == xb && ys == yb) && allowed_at(L, xs, ys) &&
                       !hist.has(PT_XY_TO_INT(xs, ys));
            }
        }
        if (pred) {
            pushed |= 1u << t;
            hist.add(PT_XY_TO_INT(xs, ys));
        }
        q_push(Q, pred, t, PT_XY_TO_INT(xs, ys));
    }
    q_run<M, TX>(Q, qbuf, L);
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
    kd = kd && !moved;
    bool kchg = false;
    if (L.kcoh_rounds) {
        const bool want = valid && !kd;
        kchg = kcoh_walk_q<M, TX>(Q, qbuf, L, want, xb, yb, db, hist);
        if (want) kd = !kchg;
    }
    bool done = was_done && !moved && !kchg;
    bool schg = false;
    const bool sp = round >= L.spatial_from;   // rounds that run the spatial window
    if (sp && L.spatial == 1) {
        schg = spatial_search_q<M, TX, true>(Q, qbuf, L, valid && !done, xb, yb, db, hist);
        if (valid && !done) done = !schg;
    } else if (sp && L.spatial > 1 && valid && !done) {
        schg = spatial_search_seq<M, TX>(L, qbuf, xb, yb, db, hist);
        done = !schg;
    } else if (L.rs && valid) {
        schg = random_search<M, TX>(L, qbuf, (long)y * L.ann_w + x, round, xb, yb, db);
    }
    if (schg) kd = false;
    if (!valid) return;
    nxt.xy[v] = PT_XY_TO_INT(xb, yb);
    nxt.d[v] = db;
    nxt.fl[v] = (done ? F_DONE : 0) | ((moved || kchg || schg) ? F_CHG : 0) | (kd ? F_KD : 0);
    if (L.hist && keep_hist) hist.store(L, v, nv);
}

// query_step > 1: every patch takes the best offset of the lattice patches within query_step
// (clamped into b, allowed positions only), starting from its initial match; a patch with no
// allowed proposal takes its lattice owner's match. Launched over all patches (qs = 1 layout).
template <int M, int TX>
__device__ __forceinline__ void gather_body(const LookupArgs &L, NnfField lat, int32_t *__restrict__ oxy,
                                            float *__restrict__ od) {
    __shared__ float4 qbuf[QBuf<M, TX>::F4];
    const int2 tc = tile_coords();
    const int x = tc.x * TX + threadIdx.x % TX, y = tc.y * (THREADS / TX) + threadIdx.x / TX;
    const bool valid = x < L.ann_w && y < L.ann_h;
    stage_query<M, TX>(L, qbuf, valid, x, y);
    if (!valid) return;
    const int bew = L.bw - 7, beh = L.bh - 7, qs = L.gqs;
    const long pi = (long)y * L.ann_w + x;
    bool have = false;
    int xb = 0, yb = 0;
    float db = 0.f;
    if (L.init_xy) {
        const int v = L.init_xy[pi];
        xb = PT_INT_TO_X(v);
        yb = PT_INT_TO_Y(v);
        db = L.init_d[pi];
        have = true;
    }
    const int vy0 = y / qs - (y % qs == 0 ? 1 : 0), vy1 = y / qs + 1;
    const int vx0 = x / qs - (x % qs == 0 ? 1 : 0), vx1 = x / qs + 1;
    for (int vy = vy0 < 0 ? 0 : vy0; vy <= vy1 && vy < L.vh; vy++)
        for (int vx = vx0 < 0 ? 0 : vx0; vx <= vx1 && vx < L.vw; vx++) {
            const int c = lat.xy[(long)vy * L.vw + vx];
            int xs = PT_INT_TO_X(c) + (x - vx * qs), ys = PT_INT_TO_Y(c) + (y - vy * qs);
            xs = xs < 0 ? 0 : (xs >= bew ? bew - 1 : xs);
            ys = ys < 0 ? 0 : (ys >= beh ? beh - 1 : ys);
            if (!allowed_at(L, xs, ys)) continue;
            if (have && xs == xb && ys == yb) continue;
            const float d = qdist<M, TX>(L, qbuf, threadIdx.x, xs, ys, have ? db : INFINITY);
            if (!have || d < db) {
                db = d;
                xb = xs;
                yb = ys;
                have = true;
            }
        }
    if (!have) {
        const int c = lat.xy[(long)(y / qs) * L.vw + x / qs];
        xb = PT_INT_TO_X(c);
        yb = PT_INT_TO_Y(c);
        db = qdist<M, TX>(L, qbuf, threadIdx.x, xb, yb, INFINITY);
    }
    oxy[pi] = PT_XY_TO_INT(xb, yb);
    od[pi] = db;
}

// MFMA form's candidate pass (image mode, query_step 1): the initial match and the cells of every
// table, scored by int8 SSD; the MFMA form walks no k-coherence lists and runs no spatial or
// random search in this pass, so none of candidates_body's queue machinery is needed.
template <int TX>
__device__ __forceinline__ void cand_int8_body(const LookupArgs &L, NnfField out) {
    __shared__ float4 qbuf[QBuf<M_INT8, TX>::F4];
    const int2 tc = tile_coords();
    const int x = tc.x * TX + threadIdx.x % TX, y = tc.y * (THREADS / TX) + threadIdx.x / TX;
    const bool valid = x < L.vw && y < L.vh;
    stage_query<M_INT8, TX>(L, qbuf, valid, x, y);
    if (!valid) return;
    const long pi = (long)y * L.ann_w + x;
    {   // |q|^2 of the lane's patch, kept for the rounds (and read back by qdist below)
        const uint32_t *qt = (const uint32_t *)qbuf + (threadIdx.x / TX) * (TX + 7) + (threadIdx.x % TX);
        int qn = 0;
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int c = 0; c < 8; c++) qn = __builtin_amdgcn_sdot4((int)qt[r * (TX + 7) + c], (int)qt[r * (TX + 7) + c], qn, false);
        const_cast<int32_t *>(L.qnorm)[pi] = qn;
    }
    bool have, moved;
    int xb, yb;
    float db;
    Hist hist;
    init_and_tables<M_INT8, TX, false>(L, qbuf, true, x, y, pi, have, moved, xb, yb, db, hist);
    out.xy[pi] = PT_XY_TO_INT(xb, yb);
    out.d[pi] = db;
    out.fl[pi] = F_CHG | (moved ? F_MOVED : 0);
}

// MFMA form's output stage (NNF_X, NNF_Y, NNF_DIST as doubles, as k_write_ann): the field holds
// int8 scores; a patch whose match is not its initial one gets the exact FP32 SSD of the match
// (the reference's lane order), a patch holding its initial (previous-field) match its initial
// distance, with the temporal rescaling of patchtable.h:4178-4193.
template <int TX>
__device__ __forceinline__ void exact_out_body(const LookupArgs &L, NnfField f, double *__restrict__ ann,
                                               const double *__restrict__ prev, double temporal_scale, int fix) {
    __shared__ float4 qbuf[QBuf<M_TILE, TX>::F4];
    const int2 tc = tile_coords();
    const int vx = tc.x * TX + threadIdx.x % TX, vy = tc.y * (THREADS / TX) + threadIdx.x / TX;
    const bool valid = vx < L.vw && vy < L.vh;
    stage_query<M_TILE, TX>(L, qbuf, valid, vx, vy);
    if (!valid) return;
    const long v = (long)vy * L.vw + vx;
    const int c = f.xy[v];
    const int x = PT_INT_TO_X(c), y = PT_INT_TO_Y(c);
    const float d = (f.fl[v] & F_MOVED) || !L.init_d ? qdist<M_TILE, TX>(L, qbuf, threadIdx.x, x, y, INFINITY)
                                                     : L.init_d[v];
    double dd = (double)d;
    if (fix && (double)x == prev[v * 3] && (double)y == prev[v * 3 + 1] && d == L.init_d[v]) dd = dd / temporal_scale;
    ann[v * 3 + 0] = (double)x;
    ann[v * 3 + 1] = (double)y;
    ann[v * 3 + 2] = dd;
}

// Kernels of one lookup (arguments by value) and of a batch of lookups of one query against many
// tables (blockIdx.z = pair; each pair's arguments and fields in a device array). A pair's lanes
// compute exactly what its single lookup computes.
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_candidates(LookupArgs L, NnfField out) {
    candidates_body<M, TX>(L, out);
}
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_propagate(LookupArgs L, int J, int same_jump, int round, NnfField cur,
                                                      NnfField nxt, int keep_hist) {
    propagate_body<M, TX>(L, J, same_jump, round, cur, nxt, keep_hist);
}
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_gather(LookupArgs L, NnfField lat, int32_t *__restrict__ oxy,
                                                    float *__restrict__ od) {
    gather_body<M, TX>(L, lat, oxy, od);
}
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_candidates_batch(const PairLookup *__restrict__ P) {
    const PairLookup &p = P[blockIdx.z];
    candidates_body<M, TX>(p.L, p.f[0]);
}
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_propagate_batch(const PairLookup *__restrict__ P, int J, int same_jump,
                                                            int round, int src, int keep_hist) {
    const PairLookup &p = P[blockIdx.z];
    propagate_body<M, TX>(p.L, J, same_jump, round, p.f[src], p.f[src ^ 1], keep_hist);
}
template <int M, int TX>
__global__ __launch_bounds__(THREADS) void k_gather_batch(const PairLookup *__restrict__ P, int src) {
    const PairLookup &p = P[blockIdx.z];
    gather_body<M, TX>(p.G, p.f[src], p.gxy, p.gd);
}
template <int TX>
__global__ __launch_bounds__(THREADS) void k_cand_int8(LookupArgs L, NnfField out) {
    cand_int8_body<TX>(L, out);
}
template <int TX>
__global__ __launch_bounds__(THREADS) void k_cand_int8_batch(const PairLookup *__restrict__ P) {
    const PairLookup &p = P[blockIdx.z];
    cand_int8_body<TX>(p.L, p.f[0]);
}
template <int TX>
__global__ __launch_bounds__(THREADS) void k_exact_out(LookupArgs L, NnfField f, double *__restrict__ ann,
                                                      const double *__restrict__ prev, double temporal_scale, int fix) {
    exact_out_body<TX>(L, f, ann, prev, temporal_scale, fix);
}
template <int TX>
__global__ __launch_bounds__(THREADS) void k_exact_out_batch(const PairLookup *__restrict__ P, int src) {
    const PairLookup &p = P[blockIdx.z];
    exact_out_body<TX>(p.L, p.f[src], p.ann, nullptr, 1.0, 0);
}

// Initial field from a previous one (TABLE_COPY_PREV_NNF and patchtable.h:3539-3585).
template <int M>
__global__ __launch_bounds__(256) void k_init_prev(LookupArgs L, const double *__restrict__ prev, const float *__restrict__ sv, int recalc,
                            double temporal_scale, int sanitize, int32_t *__restrict__ ixy, float *__restrict__ id,
                            int *__restrict__ bad) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)L.ann_w * L.ann_h) return;
    const int x = (int)(i % L.ann_w), y = (int)(i / L.ann_w);
    const int bew = L.bw - 7, beh = L.bh - 7;
    int xs = (int)prev[i * 3], ys = (int)prev[i * 3 + 1];
    if (!sanitize && (xs < 0 || ys < 0 || xs >= bew || ys >= beh)) atomicOr(bad, 1);
    xs = xs < 0 ? 0 : (xs >= bew ? bew - 1 : xs);
    ys = ys < 0 ? 0 : (ys >= beh ? beh - 1 : ys);
    ixy[i] = PT_XY_TO_INT(xs, ys);
    if (recalc) {
        float d;
        if constexpr (M == M_DESC)
            d = pt_desc_dist(L.a_descI + ((size_t)(y + 7) * (L.aw + 7) + x + 7) * L.C,
                          L.b_descI + ((size_t)(ys + 7) * (L.bw + 7) + xs + 7) * L.C, L.C);
        else
            d = pt_patch_ssd(L.a + ((size_t)y * L.aw + x) * 3, L.aw * 3, L.b + ((size_t)ys * L.bw + xs) * 3, L.bw * 3);
        const double tsc = sv ? (double)sv[i] : temporal_scale;
        id[i] = (float)(d * tsc);
    } else {
        id[i] = (float)(prev[i * 3 + 2] * temporal_scale);
    }
}

// NNF_X, NNF_Y, NNF_DIST as doubles; a patch that kept its initial match and distance gets the
// distance divided by the temporal scale (patchtable.h:4178-4193).
__global__ void k_write_ann(const int32_t *__restrict__ xy, const float *__restrict__ d, long n, double *__restrict__ ann,
                            const double *__restrict__ prev, const float *__restrict__ init_d, double temporal_scale,
                            int fix) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = xy[i];
    const int x = PT_INT_TO_X(v), y = PT_INT_TO_Y(v);
    double dd = (double)d[i];
    if (fix && (double)x == prev[i * 3] && (double)y == prev[i * 3 + 1] && d[i] == init_d[i]) dd = dd / temporal_scale;
    ann[i * 3 + 0] = (double)x;
    ann[i * 3 + 1] = (double)y;
    ann[i * 3 + 2] = dd;
}

// NNF output of every pair of a batch (no previous field): blockIdx.y = pair; field `src`, or the
// gathered field with query_step > 1.
__global__ void k_write_ann_batch(const PairLookup *__restrict__ P, int src, int gathered, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PairLookup &p = P[blockIdx.y];
    const int v = gathered ? p.gxy[i] : p.f[src].xy[i];
    const float d = gathered ? p.gd[i] : p.f[src].d[i];
    p.ann[i * 3 + 0] = (double)PT_INT_TO_X(v);
    p.ann[i * 3 + 1] = (double)PT_INT_TO_Y(v);
    p.ann[i * 3 + 2] = (double)d;
}

// Descriptor-mode input: an h_in x w_in x C descriptor image placed at rows/columns 7.. of the
// padded (Hp x Wp) descriptor (get_padded_descriptor, patchtable.h:2498-2507), or a padded input
// copied as is; written both as planes (quantiser, table cells) and interleaved (distances).
__global__ void k_desc_pad(const float *__restrict__ in, int h_in, int w_in, int C, int padded, int Hp, int Wp,
                           float *__restrict__ planes, float *__restrict__ inter) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Hp * Wp) return;
    const int y = (int)(i / Wp), x = (int)(i % Wp);
    const int sy = padded ? y : y - 7, sx = padded ? x : x - 7;
    const bool in_img = sy >= 0 && sx >= 0 && sy < h_in && sx < w_in;
    const size_t plane = (size_t)Hp * Wp;
    for (int c = 0; c < C; c++) {
        const float v = in_img ? in[((size_t)sy * w_in + sx) * C + c] : 0.f;
        planes[c * plane + i] = v;
        inter[i * C + c] = v;
    }
}

__global__ void k_ssd_pairs(const float *a, int aw, const float *b, int bw, const int32_t *__restrict__ ax,
                            const int32_t *__restrict__ ay, const int32_t *__restrict__ bx, const int32_t *__restrict__ by,
                            int n, float *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = pt_patch_ssd(a + ((size_t)ay[i] * aw + ax[i]) * 3, aw * 3, b + ((size_t)by[i] * bw + bx[i]) * 3, bw * 3);
}

__global__ void k_desc_dist_pairs(const float *a, const float *b, int C, int n, float *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = pt_desc_dist(a + (size_t)i * C, b + (size_t)i * C, C);
}

template <int M, int TX>
int launch_lookup(const LookupArgs &L, NnfField f0, NnfField f1, hipStream_t s, hipEvent_t ev_after_candidates,
                  int &launches) {
    dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX));
    hipLaunchKernelGGL((k_candidates<M, TX>), grid, dim3(THREADS), 0, s, L, f0);
    if (ev_after_candidates) PT_HIP_TRY(hipEventRecord(ev_after_candidates, s));
    const int R = L.rounds;
    for (int r = 0; r < R; r++) {
        const int J = pt_jump_at(R, r);
        const int same = r > 0 && pt_jump_at(R, r - 1) == J;
        const int keep = r + 1 < R;   // the last round's history is not read again
        hipLaunchKernelGGL((k_propagate<M, TX>), grid, dim3(THREADS), 0, s, L, J, same, r + 1, f0, f1, keep);
        std::swap(f0, f1);
    }
    launches = R;
    PT_LAUNCH_CHECK();
    return R % 2 ? 1 : 0;   // which field holds the result
}

template <int M, int TX>
void launch_candidates(const LookupArgs &L, NnfField f0, hipStream_t s) {
    dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX));
    hipLaunchKernelGGL((k_candidates<M, TX>), grid, dim3(THREADS), 0, s, L, f0);
}

// Phases of a batch: the candidates, the rounds and the gather of every pair, one launch each.
template <int M, int TX>
int launch_batch(const PairLookup *P, const PairLookup &P0, int npairs, hipStream_t s, int &src, int &launches) {
    const LookupArgs &L = P0.L;
    dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX), npairs);
    hipLaunchKernelGGL((k_candidates_batch<M, TX>), grid, dim3(THREADS), 0, s, P);
    launches++;
    const int R = L.rounds;
    src = 0;
    for (int r = 0; r < R; r++) {
        const int J = pt_jump_at(R, r);
        const int same = r > 0 && pt_jump_at(R, r - 1) == J;
        const int keep = r + 1 < R;
        hipLaunchKernelGGL((k_propagate_batch<M, TX>), grid, dim3(THREADS), 0, s, P, J, same, r + 1, src, keep);
        src ^= 1;
        launches++;
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

template <int M, int TX>
void launch_gather(const PairLookup *P, const LookupArgs &G, int npairs, int src, hipStream_t s) {
    dim3 grid((G.ann_w + TX - 1) / TX, (G.ann_h + THREADS / TX - 1) / (THREADS / TX), npairs);
    hipLaunchKernelGGL((k_gather_batch<M, TX>), grid, dim3(THREADS), 0, s, P, src);
}

template <int M, int TX>
void launch_gather1(const LookupArgs &G, NnfField res, int32_t *gxy, float *gd, hipStream_t s) {
    dim3 grid((G.ann_w + TX - 1) / TX, (G.ann_h + THREADS / TX - 1) / (THREADS / TX));
    hipLaunchKernelGGL((k_gather<M, TX>), grid, dim3(THREADS), 0, s, G, res, gxy, gd);
}

// Calls f(std::integral_constant<int, TX>) for the tile width `tile_x` (64, 32 or 16; descriptor
// mode always 64: its LDS holds one descriptor per lane whatever the shape).
template <class F>
int with_tile(int tile_x, int C, F &&f) {
    if (C > 0 || tile_x == 64) return f(std::integral_constant<int, 64>{});
    if (tile_x == 32) return f(std::integral_constant<int, 32>{});
    if (tile_x == 16) return f(std::integral_constant<int, 16>{});
    return pt_failf(PT_E_INVALID, "tile width %d is not one of 64, 32, 16", tile_x);
}

}  // namespace

int pt_lookup_tile_ok(int tile_x) { return tile_x == 64 || tile_x == 32 || tile_x == 16; }

int pt_lookup_init_prev(const LookupArgs &L, const double *prev, const float *sv, int recalc, double temporal_scale,
                        int sanitize, int32_t *ixy, float *id, int *bad_dev, hipStream_t s) {
    const long n = (long)L.ann_w * L.ann_h;
    if (L.C > 0)
        hipLaunchKernelGGL(k_init_prev<M_DESC>, dim3(pt_blocks(n, 256)), dim3(256), 0, s, L, prev, sv, recalc,
                           temporal_scale, sanitize, ixy, id, bad_dev);
    else
        hipLaunchKernelGGL(k_init_prev<M_GLOBAL>, dim3(pt_blocks(n, 256)), dim3(256), 0, s, L, prev, sv, recalc,
                           temporal_scale, sanitize, ixy, id, bad_dev);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_lookup_candidates(const LookupArgs &L, NnfField f0, hipStream_t s) {
    if (L.C > PT_LOOKUP_CMAX) return pt_failf(PT_E_UNSUPPORTED, "descriptor dims %d > %d", L.C, PT_LOOKUP_CMAX);
    return with_tile(L.tile_x, L.C, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        if (L.C > 0) {
            if constexpr (TX == 64) launch_candidates<M_DESC, TX>(L, f0, s);
        } else if (L.int8) {
            dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX));
            hipLaunchKernelGGL((k_cand_int8<TX>), grid, dim3(THREADS), 0, s, L, f0);
        } else if (L.qs == 1) launch_candidates<M_TILE, TX>(L, f0, s);
        else launch_candidates<M_GLOBAL, TX>(L, f0, s);
        PT_LAUNCH_CHECK();
        return PT_OK;
    });
}

int pt_lookup_output(const LookupArgs &L, NnfField res, int32_t *gxy, float *gd, double *ann_out, const double *prev,
                     const float *init_d, double temporal_scale, int temporal_fix, hipStream_t s, int *launches,
                     hipEvent_t ev_before_write) {
    const long n = (long)L.ann_w * L.ann_h;
    const int32_t *oxy = res.xy;
    const float *od = res.d;
    if (L.qs > 1) {
        LookupArgs G = L;
        G.gqs = L.qs;
        G.qs = 1;   // k_gather lanes are single patches; stage_query / qdist use qs 1
        TRY_PT(with_tile(L.tile_x, L.C, [&](auto tx_) {
            constexpr int TX = decltype(tx_)::value;
            if (L.C > 0) {
                if constexpr (TX == 64) launch_gather1<M_DESC, TX>(G, res, gxy, gd, s);
            } else {
                launch_gather1<M_TILE, TX>(G, res, gxy, gd, s);
            }
            return PT_OK;
        }));
        oxy = gxy;
        od = gd;
        if (launches) (*launches)++;
    }
    if (ev_before_write) PT_HIP_TRY(hipEventRecord(ev_before_write, s));
    hipLaunchKernelGGL(k_write_ann, dim3(pt_blocks(n, 256)), dim3(256), 0, s, oxy, od, n, ann_out, prev, init_d,
                       temporal_scale, temporal_fix);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_lookup_launch(const LookupArgs &L, NnfField f0, NnfField f1, int32_t *gxy, float *gd, double *ann_out,
                     const double *prev, const float *init_d, double temporal_scale, int temporal_fix, hipStream_t s,
                     hipEvent_t ev_after_candidates, hipEvent_t ev_after_rounds, int *launches) {
    if (L.C > PT_LOOKUP_CMAX) return pt_failf(PT_E_UNSUPPORTED, "descriptor dims %d > %d", L.C, PT_LOOKUP_CMAX);
    int which = 0, nl = 0;
    TRY_PT(with_tile(L.tile_x, L.C, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        if (L.C > 0) {
            if constexpr (TX == 64) which = launch_lookup<M_DESC, TX>(L, f0, f1, s, ev_after_candidates, nl);
        } else if (L.qs == 1) which = launch_lookup<M_TILE, TX>(L, f0, f1, s, ev_after_candidates, nl);
        else which = launch_lookup<M_GLOBAL, TX>(L, f0, f1, s, ev_after_candidates, nl);
        return which < 0 ? which : PT_OK;
    }));
    NnfField res = which ? f1 : f0;
    const int rc = pt_lookup_output(L, res, gxy, gd, ann_out, prev, init_d, temporal_scale, temporal_fix, s, &nl,
                                    ev_after_rounds);
    if (rc != PT_OK) return rc;
    if (launches) *launches = nl;
    return PT_OK;
}

int pt_lookup_batch_launch(const PairLookup *P_dev, const PairLookup &P0, int npairs, hipStream_t s,
                           hipEvent_t ev_after_rounds, int *launches) {
    const LookupArgs &L = P0.L;
    if (L.C > 0) return pt_fail(PT_E_UNSUPPORTED, "batched lookups take images, not descriptors");
    if (npairs < 1 || npairs > 65535) return pt_failf(PT_E_INVALID, "batch of %d pairs", npairs);
    int src = 0, nl = 0;
    TRY_PT(with_tile(L.tile_x, 0, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        if (L.qs == 1) return launch_batch<M_TILE, TX>(P_dev, P0, npairs, s, src, nl);
        return launch_batch<M_GLOBAL, TX>(P_dev, P0, npairs, s, src, nl);
    }));
    TRY_PT(pt_lookup_batch_output(P_dev, P0, npairs, src, s, ev_after_rounds, &nl));
    if (launches) *launches = nl;
    return PT_OK;
}

int pt_lookup_batch_candidates(const PairLookup *P_dev, const PairLookup &P0, int npairs, hipStream_t s) {
    const LookupArgs &L = P0.L;
    if (L.C > 0) return pt_fail(PT_E_UNSUPPORTED, "batched lookups take images, not descriptors");
    return with_tile(L.tile_x, 0, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX), npairs);
        if (L.int8) hipLaunchKernelGGL((k_cand_int8_batch<TX>), grid, dim3(THREADS), 0, s, P_dev);
        else if (L.qs == 1) hipLaunchKernelGGL((k_candidates_batch<M_TILE, TX>), grid, dim3(THREADS), 0, s, P_dev);
        else hipLaunchKernelGGL((k_candidates_batch<M_GLOBAL, TX>), grid, dim3(THREADS), 0, s, P_dev);
        PT_LAUNCH_CHECK();
        return PT_OK;
    });
}

int pt_lookup_exact_out(const LookupArgs &L, NnfField f, double *ann_out, const double *prev, double temporal_scale,
                        int temporal_fix, hipStream_t s) {
    if (L.C > 0 || L.qs != 1) return pt_fail(PT_E_UNSUPPORTED, "int8 scores are kept in image mode with query_step 1");
    return with_tile(L.tile_x, 0, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX));
        hipLaunchKernelGGL((k_exact_out<TX>), grid, dim3(THREADS), 0, s, L, f, ann_out, prev, temporal_scale,
                           temporal_fix);
        PT_LAUNCH_CHECK();
        return PT_OK;
    });
}

int pt_lookup_batch_exact_out(const PairLookup *P_dev, const PairLookup &P0, int npairs, int src, hipStream_t s) {
    const LookupArgs &L = P0.L;
    if (L.C > 0 || L.qs != 1) return pt_fail(PT_E_UNSUPPORTED, "int8 scores are kept in image mode with query_step 1");
    return with_tile(L.tile_x, 0, [&](auto tx_) {
        constexpr int TX = decltype(tx_)::value;
        dim3 grid((L.vw + TX - 1) / TX, (L.vh + THREADS / TX - 1) / (THREADS / TX), npairs);
        hipLaunchKernelGGL((k_exact_out_batch<TX>), grid, dim3(THREADS), 0, s, P_dev, src);
        PT_LAUNCH_CHECK();
        return PT_OK;
    });
}

int pt_lookup_batch_output(const PairLookup *P_dev, const PairLookup &P0, int npairs, int src, hipStream_t s,
                           hipEvent_t ev_before_write, int *launches) {
    const LookupArgs &L = P0.L;
    const int gathered = L.qs > 1;
    if (gathered) {
        TRY_PT(with_tile(L.tile_x, 0, [&](auto tx_) {
            constexpr int TX = decltype(tx_)::value;
            launch_gather<M_TILE, TX>(P_dev, P0.G, npairs, src, s);
            return PT_OK;
        }));
        if (launches) (*launches)++;
    }
    if (ev_before_write) PT_HIP_TRY(hipEventRecord(ev_before_write, s));
    const long n = (long)L.ann_w * L.ann_h;
    hipLaunchKernelGGL(k_write_ann_batch, dim3(pt_blocks(n, 256), npairs), dim3(256), 0, s, P_dev, src, gathered, n);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_desc_pad_launch(const float *in, int h_in, int w_in, int C, int padded, int Hp, int Wp, float *planes,
                       float *inter, hipStream_t s) {
    const long n = (long)Hp * Wp;
    hipLaunchKernelGGL(k_desc_pad, dim3(pt_blocks(n, 256)), dim3(256), 0, s, in, h_in, w_in, C, padded, Hp, Wp, planes,
                       inter);
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

int pt_desc_dist_pairs_launch(const float *a, const float *b, int C, int n, float *out, hipStream_t s) {
    if (n <= 0) return PT_OK;
    hipLaunchKernelGGL(k_desc_dist_pairs, dim3(pt_blocks(n, 256)), dim3(256), 0, s, a, b, C, n, out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
