// MFMA lookup rounds on gfx950: dense int8 candidate scoring on matrix cores.
//
// The plain lookup (lookup.hip) evaluates each candidate with one exact FP32 SSD per lane: 8
// rows x 96 bytes gathered at an arbitrary place in b, about 48 tag lookups of L1, and the
// lookup is bound by that gather rate. Here the candidates of 16 query patches that share a
// nearby offset are scored together as a dense contraction on v_mfma_i32_16x16x64_i8:
//
//   for query q and database patch p (8x8 RGB quantised to int8: round(255 clamp(v, 0, 1)) - 128,
//   laid out RGB0, so a patch is a 256-byte vector with 64 zero bytes),
//       |q - p|^2 = |q|^2 + |p|^2 - 2 q.p ;
//   an exact int32 (the MFMA accumulates int8 products in int32: no rounding anywhere).
//
// The field of the MFMA form holds each patch's match and the match's int8 SSD (its score; the
// candidate pass scores the table cells the same way, lookup.hip M_INT8). One round with jump J,
// for each group of 16 consecutive query patches of a row (x0 = 16k):
//   sources: 64 positions -- the 16 members, the 16 patches J rows above and J rows below,
//     8 patches J to the left of members 0-7 and 8 patches J to the right of members 8-15;
//     each valid source proposes its current offset (match - position);
//   anchors: the distinct offsets in slot order of first occurrence with their counts; up to
//     mfma_anchors times take the uncovered offset of largest count (ties: first slot) and mark
//     every offset within 8 columns and R rows of it covered;
//   search: for anchor (ox, oy), database patches x0 + ox - 8 + j (j < 32) on rows
//     y + oy + dy (|dy| <= R) -- each member sees every offset within 8 columns of the anchor;
//     each member keeps the candidate of smallest (SSD, anchor, dy, j), SSD = |q|^2 + |p|^2 - 2 q.p
//     the exact integer SSD of the int8 patches;
//   acceptance: a member whose best candidate differs from its match takes it when its int8 SSD
//     is strictly below the match's score.
// After the rounds the exact FP32 SSD of every final match (the reference's lane order,
// patch_dist_exact, patchtable.h:330-373) replaces its score (lookup.hip), so NNF_DIST is always
// the exact distance of the stored match. The oracle (lookup_parallel with mfma_rounds > 0)
// restates this bit for bit: the integer scores do not depend on summation order.
//
// Work layout: a workgroup of one wave owns 64 query patches of a row, its 4 groups in turn. A
// wave stages each anchor's band of b (39 x 10 pixels, int8 RGB0) in its own LDS region; each
// lane loads the six B fragments of one candidate column that the band's 3 rows of candidates
// share, and the 3 x 4 MFMAs of a 16-candidate block read them from registers. After the first
// round a group does not search again an anchor it searched in its last round with anchors,
// when no score exceeds the int8 SSD of its match (see mfma_round_body): the decisions are
// those of the full search.
#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_tile.h"

namespace {

using namespace pt_tile;

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int R = PT_MFMA_BAND;        // rows of candidates above and below the anchor
constexpr int BR = 8 + 2 * R;          // band rows
constexpr int BC = 32 + 7;             // band columns (32 candidates of 8 pixels)
constexpr uint64_t NONE = ~0ull;
constexpr uint8_t F_CHG = 2;           // field flag: the round that wrote the field changed the match
constexpr uint8_t F_MOVED = PT_FIELD_MOVED;   // ... the match is not the initial one (int8 score)
// No packed offset equals this: both halves would be -32768, and |offset| <= 32767.
constexpr int HT_EMPTY = (int)0x80008000u;

__device__ __forceinline__ uint32_t quant_pixel(const float *p) {
    uint32_t u = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float v = fminf(fmaxf(p[c], 0.f), 1.f) * 255.f;
        const int q = (int)rintf(v) - 128;
        u |= ((uint32_t)q & 0xffu) << (8 * c);
    }
    return u;
}

// int8 RGB0 copy of an RGB float image.
__global__ void k_quant_rgb0(const float *__restrict__ img, long npx, uint32_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npx) out[i] = quant_pixel(img + i * 3);
}

__device__ __forceinline__ int sq_pixel(uint32_t u) {
    const int r = (int)(int8_t)(u & 0xff), g = (int)(int8_t)((u >> 8) & 0xff), b = (int)(int8_t)((u >> 16) & 0xff);
    return r * r + g * g + b * b;
}

constexpr int P = PT_MFMA_PAD;

// int8 RGB0 copy of an h x w RGB image into the interior of a (h + 2P) x (w + 2P) array.
__global__ void k_quant_rgb0_pad(const float *__restrict__ img, int h, int w, uint32_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)h * w) return;
    const int x = (int)(i % w), y = (int)(i / w);
    out[(size_t)(y + P) * (w + 2 * P) + x + P] = quant_pixel(img + i * 3);
}

// Norm key of every 8x8 patch of the padded int8 copy (pw x ph patches), (|p|^2 + OFF) << 7
// (see phase 2 of the round), into the interior of a (ph + 2P) x (pw + 2P) array whose margins
// hold PT_MFMA_NK_PAD.
__global__ void k_patch_norms_pad(const uint32_t *__restrict__ q, int w, int pw, int ph, uint32_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)pw * ph) return;
    const int x = (int)(i % pw), y = (int)(i / pw);
    int s = 0;
    for (int r = 0; r < 8; r++) {
        const uint32_t *row = q + (size_t)(y + r + P) * (w + 2 * P) + x + P;
#pragma unroll
        for (int c = 0; c < 8; c++) s += sq_pixel(row[c]);
    }
    out[(size_t)(y + P) * (pw + 2 * P) + x + P] = ((uint32_t)s + PT_MFMA_OFF) << 7;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const int lo = __shfl((int)(uint32_t)v, src), hi = __shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m), hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
// Minimum over each row of 16 lanes, left in every lane of the row: DPP swaps within quads, then
// the half-row mirror (lane i <-> 7 - i) joins the two quads of each half, the row mirror
// (i <-> 15 - i) the two halves.
__device__ __forceinline__ uint32_t row16_min(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true));   // quad_perm 1,0,3,2
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true));   // quad_perm 2,3,0,1
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true));  // row_half_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true));  // row_mirror
    return v;
}
__device__ __forceinline__ uint32_t row16_max(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true));
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true));
    return v;
}
// The 64-bit row minimum, by the same four DPP steps on both halves.
#define PT_DPP64_STEP(v, ctrl)                                                                           \
    {                                                                                                    \
        const uint32_t lo_ = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(v), ctrl, 0xF, 0xF, true); \
        const uint32_t hi_ = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)((v) >> 32), ctrl, 0xF, 0xF, true); \
        const uint64_t o_ = ((uint64_t)hi_ << 32) | lo_;                                                 \
        if (o_ < (v)) (v) = o_;                                                                          \
    }
__device__ __forceinline__ uint64_t row16_min64(uint64_t v) {
    PT_DPP64_STEP(v, 0xB1)
    PT_DPP64_STEP(v, 0x4E)
    PT_DPP64_STEP(v, 0x141)
    PT_DPP64_STEP(v, 0x140)
    return v;
}
#undef PT_DPP64_STEP
// Orders this wave's LDS writes before its later reads (and reads before later writes).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

constexpr int NB = 32 * (2 * R + 1);               // candidate norms of one band
constexpr int BANDW = BR * BC + NB;                // words of one band buffer
constexpr int MAXITEMS = 4 * PT_MFMA_MAX_ANCHORS;  // (group, anchor) items of one wave

constexpr int PER = (BANDW + 63) / 64;   // band words per lane
constexpr int HS = 96;                   // hash slots per group (64 sources)
// phase 1 needs 4 groups x (3 HS table words + 64 source offsets) + the anchors
constexpr int WSH = 12 * HS + 256 + 8 * PT_MFMA_MAX_ANCHORS > 2 * BANDW ? 12 * HS + 256 + 8 * PT_MFMA_MAX_ANCHORS
                                                                        : 2 * BANDW;

// Per-lane byte offsets of the band words this lane loads: pixel (r, c) of the band at
// r * bwp + c in the padded copy of b, or norm key (dy, j) at dy * bnp + j in the padded norm key
// table (-1: none).
// They do not depend on the item, so they are computed once.
__device__ __forceinline__ void band_offsets(const MfmaArgs &M, int lane, int (&off)[PER]) {
    const int bwp = M.bw + 2 * P, bnp = M.bw - 7 + 2 * P;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int e = lane + 64 * k;
        if (e < BR * BC) {
            const int r = (e * 1681) >> 16, c = e - r * BC;   // e / 39 for e < 1000
            off[k] = (r * bwp + c) * 4;
        } else if (e < BANDW) {
            const int n = e - BR * BC;
            off[k] = ((n >> 5) * bnp + (n & 31)) * 4;
        } else {
            off[k] = -1;
        }
    }
}

// Loads the band of item (x0, ox, oy) into registers from the padded copies, as buffer loads
// whose base is the band's origin (scalar) and whose offsets are the lane's fixed band offsets.
// The band origin lies inside the margins (see PT_MFMA_PAD); it is clamped to them anyway, and
// each resource ends where its array ends, so a load can never leave the arrays.
__device__ __forceinline__ void band_fetch(const MfmaArgs &M, int x0, int y, int ox, int oy, const int (&off)[PER],
                                           uint32_t (&reg)[PER]) {
    const int bwp = M.bw + 2 * P, bnp = M.bw - 7 + 2 * P;
    const int u0 = min(max(x0 + ox - 8, -P), M.bw + P - BC), v0 = min(max(y + oy - R, -P), M.bh + P - BR);
    const size_t oq = (size_t)(v0 + P) * bwp + (u0 + P), on = (size_t)(v0 + P) * bnp + (u0 + P);
    const size_t nq = (size_t)(M.bh + 2 * P) * bwp, nn = (size_t)(M.bh - 7 + 2 * P) * bnp;
    const __amdgpu_buffer_rsrc_t rq = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(M.bq + oq), 0, (int)min((nq - oq) * 4, (size_t)0x7fffffff), 0x00020000);
    const __amdgpu_buffer_rsrc_t rn = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(M.bnorm + on), 0, (int)min((nn - on) * 4, (size_t)0x7fffffff), 0x00020000);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int e = threadIdx.x + 64 * k;
        if (64 * k + 63 < BR * BC) {               // pixels in every lane
            reg[k] = __builtin_amdgcn_raw_buffer_load_b32(rq, off[k], 0, 0);
        } else if (64 * k < BR * BC) {             // pixels, then norm keys: both loads in every
            const bool px = e < BR * BC;           // lane (a per-lane choice of resource would
            const uint32_t vq = __builtin_amdgcn_raw_buffer_load_b32(rq, px ? off[k] : 0, 0, 0);   // loop)
            const uint32_t vn = __builtin_amdgcn_raw_buffer_load_b32(rn, px ? 0 : off[k], 0, 0);
            reg[k] = px ? vq : vn;
        } else if (64 * k < BANDW) {               // norm keys (lanes past the band load word 0)
            const uint32_t vn = __builtin_amdgcn_raw_buffer_load_b32(rn, e < BANDW ? off[k] : 0, 0, 0);
            reg[k] = e < BANDW ? vn : 0u;
        } else {
            reg[k] = 0u;
        }
    }
}

// One wave per workgroup: the workgroup owns 64 x 1 query patches, so a wave leaves as soon as its
// own items are done (its row's item count varies with the image).
constexpr int MW = 64;   // threads

// `stable`: the previous round had the same jump J and its field's flags mark the patches whose
// match it changed. A group none of whose valid sources changed would then compute the anchors,
// winners and rescorings of the previous round again and keep every match: it takes no items.
// `seen`: every match's score is at most the int8 SSD of the match, so after a round each
// member's score is at most that of every candidate the round scored for it, and scores only
// fall. An anchor the group searched in its last searched round (M.anch) then holds no candidate
// strictly below any member's score: the round keeps its place in the anchor order (the order
// bits of the other anchors' keys) but scores none of its candidates. A winner strictly below a
// member's score lies outside every such band, so the decisions are those of the full search.
__device__ __forceinline__ void mfma_round_body(const MfmaArgs &M, NnfField cur, NnfField nxt, int J, int stable,
                                                int seen) {
    // the hash tables, source offsets and anchor indices of phase 1, then the two band buffers of phase 2
    __shared__ uint32_t wsh[1][WSH];
    __shared__ int items[1][MAXITEMS];                 // group | anchor index << 2
    __shared__ int anch[4 * PT_MFMA_MAX_ANCHORS];      // the 4 groups' anchors (packed offsets)
    __shared__ uint32_t fin[16 * 2];      // (key, jj) of the 16 members at the end of a group
    const int2 tc = tile_coords();
    const int x00 = tc.x * TILE_X, y00 = tc.y;
    const int wave = 0, lane = threadIdx.x;
    const int g = lane >> 4, jj = lane & 15;
    const int y = y00 + wave, xw0 = x00;
    const int aw = M.ann_w, ah = M.ann_h;
    uint64_t mykey = NONE;   // best candidate of query (xw0 + lane, y)
    uint32_t mys = 0;        // ... its key without the order bits: |p|^2 + OFF - 2 q.p
    int nitems = 0;
    if (y < ah) {
        // ---- phase 1: sources, distinct offsets and anchors of the wave's 4 groups, all four at
        // once: row gq = lane >> 4 of the wave serves group gq, each lane 4 of its 64 source slots
        // (slot 16 r + c in register r, c = lane & 15). An offset is packed into one int (x in
        // the low 16 bits): widths and heights are at most 32768.
        int *hk = (int *)wsh[wave], *hc = hk + 4 * HS, *hf = hk + 8 * HS, *pos = hk + 12 * HS, *itmp = pos + 256;
        const int gq = lane >> 4, c = lane & 15, x0 = xw0 + 16 * gq;
        // first anchor word of the group
        auto gbase = [&]() { return ((long)y * ((aw + 15) >> 4) + (x0 >> 4)) * M.nanchors; };
        int sxy[4], sx[4], sy[4], ox[4], oy[4], h[4];
        bool sv[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            // slots 0-15 members, 16-31 J rows above, 32-47 J rows below, 48-55 J to the left of
            // members 0-7, 56-63 J to the right of members 8-15
            sx[r] = r < 3 ? x0 + c : (c < 8 ? x0 + c - J : x0 + c + J);
            sy[r] = r == 1 ? y - J : (r == 2 ? y + J : y);
            sv[r] = x0 < aw && sx[r] >= 0 && sy[r] >= 0 && sx[r] < aw && sy[r] < ah;
            sxy[r] = sv[r] ? cur.xy[(long)sy[r] * aw + sx[r]] : 0;   // all four loads in flight
        }
        // a stable round in which no source of the wave's four groups moved: the wave has no items
        bool idle = false;
        if (stable) {
            bool moved = false;
#pragma unroll
            for (int r = 0; r < 4; r++) moved = moved || (sv[r] && (cur.fl[(long)sy[r] * aw + sx[r]] & F_CHG));
            const uint64_t mv = __ballot(moved);
            if (((mv >> (16 * gq)) & 0xffffull) == 0)   // no source of this group moved
#pragma unroll
                for (int r = 0; r < 4; r++) sv[r] = false;
            idle = mv == 0;
        }
        if (!idle) {
            for (int i = lane; i < 4 * HS; i += 64) {
                hk[i] = HT_EMPTY;
                hc[i] = 0;
                hf[i] = 64;
            }
            wave_lds_sync();
            // count and first slot of each source's offset: one hash table per group in LDS (open
            // addressing, HS > 64 slots, so probing ends); counts and first slots do not depend on
            // the order in which the atomics land
#pragma unroll
            for (int r = 0; r < 4; r++) {
                ox[r] = PT_INT_TO_X(sxy[r]) - sx[r];
                oy[r] = PT_INT_TO_Y(sxy[r]) - sy[r];
                const int po = sv[r] ? (int)(((uint32_t)ox[r] & 0xffffu) | ((uint32_t)oy[r] << 16)) : 0;
                pos[64 * gq + 16 * r + c] = po;
                int hh = gq * HS + (int)__umulhi((uint32_t)po * 2654435761u, (uint32_t)HS);
                if (sv[r]) {
                    for (;;) {
                        const int prev = atomicCAS(hk + hh, HT_EMPTY, po);
                        if (prev == HT_EMPTY || prev == po) break;
                        hh = hh + 1 == (gq + 1) * HS ? gq * HS : hh + 1;
                    }
                    atomicAdd(hc + hh, 1);
                    atomicMin(hf + hh, 16 * r + c);
                }
                h[r] = hh;
            }
            wave_lds_sync();
            // anchor key of a source that is its offset's first occurrence: (count, -slot) + 1
            uint32_t kb[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int slot = 16 * r + c;
                kb[r] = sv[r] && hf[h[r]] == slot ? (uint32_t)((hc[h[r]] << 6) | (63 - slot)) + 1u : 0u;
            }
            // the group's anchors of its last searched round, one per lane of the row
            const int pv = seen && x0 < aw && c < M.nanchors ? M.anch[gbase() + c] : PT_MFMA_NO_ANCHOR;
            bool cov[4] = {false, false, false, false};
            int na = 0;   // anchors of group gq so far (the same in all lanes of the row)
            int ns = 0;   // ... of them to search
            for (int a = 0; a < M.nanchors; a++) {
                // per group, the uncovered offset of largest count, ties to the first slot (keys are unique)
                uint32_t k = 0u;
#pragma unroll
                for (int r = 0; r < 4; r++) k = max(k, cov[r] ? 0u : kb[r]);
                k = row16_max(k);
                if (!__any(k != 0u)) break;
                if (k != 0u) {
                    const int pa = pos[64 * gq + 63 - (int)((k - 1u) & 63u)];
                    const int ax = (int)(int16_t)(pa & 0xffff), ay = pa >> 16;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        cov[r] = cov[r] || ((uint32_t)(ox[r] - ax + 8) <= 16u && (uint32_t)(oy[r] - ay + R) <= (uint32_t)(2 * R));
                    // searched in the group's last searched round (rows whose k is 0 are inactive here)
                    const bool was = ((__ballot(pv == pa) >> (16 * gq)) & 0xffffull) != 0;
                    if (c == 0) {
                        anch[gq * PT_MFMA_MAX_ANCHORS + a] = pa;
                        if (!was) itmp[gq * PT_MFMA_MAX_ANCHORS + ns] = a;
                    }
                    ns += was ? 0 : 1;
                    na++;
                }
            }
            wave_lds_sync();
            // this round's anchors for the next rounds (a group that found none, i.e. a stable one,
            // keeps the anchors it last searched)
            if (M.anch && x0 < aw && na > 0 && c < M.nanchors)
                M.anch[gbase() + c] = c < na ? anch[gq * PT_MFMA_MAX_ANCHORS + c] : PT_MFMA_NO_ANCHOR;
            // the item list: group 0's anchors to search in order, then group 1's, ...
            const int n0 = __builtin_amdgcn_readlane(ns, 0), n1 = __builtin_amdgcn_readlane(ns, 16),
                      n2 = __builtin_amdgcn_readlane(ns, 32), n3 = __builtin_amdgcn_readlane(ns, 48);
            const int pre = gq == 0 ? 0 : (gq == 1 ? n0 : (gq == 2 ? n0 + n1 : n0 + n1 + n2));
            if (c < ns) items[wave][pre + c] = gq | (itmp[gq * PT_MFMA_MAX_ANCHORS + c] << 2);
            nitems = n0 + n1 + n2 + n3;
        }
    }
    if (M.items_out && lane == 0) M.items_out[(long)blockIdx.y * gridDim.x + blockIdx.x] = nitems;
    wave_lds_sync();
    // ---- phase 2: the items in order, the next item's band loading during this one's MFMAs.
    // Each lane scores one candidate column jj against the 4 members 4g + t. Within a lane a
    // candidate's key is (|p|^2 + OFF - 2 q.p) << 7 | ord with ord = (anchor * (2R+1) + dy + R) * 2
    // + block: |q|^2 is the same for all candidates of a member, and OFF >= |q|^2 makes the key
    // 128 (SSD - |q|^2 + OFF) + ord a non-negative int below 2^31 that ranks by (SSD, ord). The
    // norm key table holds (|p|^2 + OFF) << 7, and PT_MFMA_NK_PAD for positions outside b, whose
    // keys then lie in [2.48e9, 4.1e9]: above every key inside b, and at or above INV_LO. At the
    // end of a group the lanes' minima are reduced as (key, jj), i.e. (SSD, anchor, dy, j) with
    // j = 16 block + jj.
    constexpr uint32_t INV_LO = 2400000000u;
    static_assert((uint64_t)PT_MFMA_NK_PAD - 256ull * 192 * 128 * 128 >= INV_LO, "pad keys stay invalid");
    static_assert((uint64_t)PT_MFMA_NK_PAD + 127 + 256ull * 192 * 128 * 128 < (1ull << 32), "pad keys do not wrap");
    static_assert(((uint64_t)4 * PT_MFMA_OFF << 7) + 127 < INV_LO, "keys inside b stay below INV_LO");
    // item fields are wave-uniform: read them into scalar registers
    auto item = [&](int i) { return __builtin_amdgcn_readfirstlane(items[wave][i]); };
    auto item_ox = [&](int e) { return (int)(int16_t)(__builtin_amdgcn_readfirstlane(anch[(e & 3) * PT_MFMA_MAX_ANCHORS + (e >> 2)]) & 0xffff); };
    auto item_oy = [&](int e) { return __builtin_amdgcn_readfirstlane(anch[(e & 3) * PT_MFMA_MAX_ANCHORS + (e >> 2)]) >> 16; };
    auto fetch = [&](int i, const int (&off)[PER], uint32_t (&reg)[PER]) {
        const int e = item(i);
        band_fetch(M, xw0 + 16 * (e & 3), y, item_ox(e), item_oy(e), off, reg);
    };
    // -256 in a scalar register the compiler does not see as a constant: a constant factor would
    // be strength-reduced to a shift and a subtract (two instructions instead of one v_mad_i32_i24)
    int m256;
    asm("s_mov_b32 %0, -256" : "=s"(m256));
    int off[PER];
    band_offsets(M, lane, off);
    // Software pipeline: at item it the registers hold the band of item it + 1 (loaded during item
    // it - 1); they go to the LDS buffer item it - 1 read, the loads of item it + 2 start, and item
    // it is scored from the buffer written during item it - 1.
    uint32_t reg[PER];
    uint32_t *bufs = wsh[wave];
    if (nitems > 0) {
        fetch(0, off, reg);
        wave_lds_sync();   // phase 1's reads of the hash table are done
#pragma unroll
        for (int k = 0; k < PER; k++)
            if (lane + 64 * k < BANDW) bufs[lane + 64 * k] = reg[k];
        if (nitems > 1) fetch(1, off, reg);
    }
    v4i A[4];
    uint32_t best[4] = {~0u, ~0u, ~0u, ~0u};
    int cur_g = -1;   // group of the current items
    const int bew = M.bw - 7, beh = M.bh - 7;
    // this lane's first band word: row 4 (g >> 1), column jj + 4 (g & 1); the fragments and norm
    // keys it reads lie at constant distances from it
    const int lofs = 4 * (g >> 1) * BC + jj + 4 * (g & 1);
    for (int it = 0; it < nitems; it++) {
        const int ei = item(it), gi = ei & 3, ga = ei >> 2;   // group, anchor index
        const int x0 = xw0 + 16 * gi;
        wave_lds_sync();   // item it - 1's reads are done, item it's band is written
        if (it + 1 < nitems) {
            uint32_t *nb1 = bufs + ((it + 1) & 1) * BANDW;
#pragma unroll
            for (int k = 0; k < PER; k++)
                if (lane + 64 * k < BANDW) nb1[lane + 64 * k] = reg[k];
            if (it + 2 < nitems) fetch(it + 2, off, reg);
        }
        const uint32_t *bnd = bufs + (it & 1) * BANDW;
        const uint32_t *bl = bnd + lofs;
        if (gi != cur_g) {
            cur_g = gi;
            // A fragments of member jj: k-step s covers patch row s + 4 (g >> 1), pixels 4(g & 1)..+3,
            // read from the int8 query (a valid member's patch lies inside it)
            const bool mv = x0 + jj < aw;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const uint32_t *src = M.aq + (size_t)(y + s + 4 * (g >> 1)) * M.aw + x0 + jj + 4 * (g & 1);
                A[s] = mv ? *(const v4i *)__builtin_assume_aligned(src, 4) : v4i{0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            // B fragment k: band row k + 4 (g >> 1), words 16 b + jj + 4 (g & 1) .. +3; candidate row
            // dyi at k-step s uses fragment dyi + s, so the three rows share six fragments
            // (fragment k is loaded just before row k - 3 first needs it and dropped after row k,
            // so at most five are held)
            v4i F[6];
            auto frag = [&](int k) {
                const uint32_t *p = bl + k * BC + 16 * b;
                F[k] = v4i{(int)p[0], (int)p[1], (int)p[2], (int)p[3]};
            };
            static_assert(R == 1, "six fragments cover candidate rows -1, 0, +1");
#pragma unroll
            for (int k = 0; k < 4; k++) frag(k);
#pragma unroll
            for (int dyi = 0; dyi < 2 * R + 1; dyi++) {
                if (dyi > 0) frag(dyi + 3);
                v4i C = v4i{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < 4; s++) C = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[s], F[dyi + s], C, 0, 0, 0);
                const uint32_t nk = bnd[BR * BC + dyi * 32 + 16 * b + jj];   // read for every lane (no branch)
                const uint32_t base = nk | (uint32_t)((ga * (2 * R + 1) + dyi) * 2 + b);
                // C[t]: member 4g + t against candidate column jj; |q.p| < 2^22 fits the signed
                // 24-bit multiply, so base - 256 q.p is one multiply-add
#pragma unroll
                for (int t = 0; t < 4; t++) best[t] = min(best[t], base + (uint32_t)__mul24(C[t], m256));
            }
        }
        if (it + 1 == nitems || (item(it + 1) & 3) != gi) {
            // best (key, jj) of member 4g + t over the group's 16 candidate lanes: the row minimum
            // of the keys, then the smallest jj holding it; row g's results go through LDS to
            // member i = 4g + t, lane 16 gi + i
            wave_lds_sync();   // the previous group's reads of fin are done
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t m = row16_min(best[t]);
                const uint32_t jm = row16_min(best[t] == m ? (uint32_t)jj : 16u);
                if (jj == 0) {
                    fin[(4 * g + t) * 2] = m;
                    fin[(4 * g + t) * 2 + 1] = jm;
                }
            }
            wave_lds_sync();
            const uint32_t key = fin[jj * 2], kj = fin[jj * 2 + 1];
            if (g == gi && key < INV_LO) {
                const int ord = (int)(key & 127u), j = 16 * (ord & 1) + (int)kj, dyi = (ord >> 1) % (2 * R + 1),
                          an = (ord >> 1) / (2 * R + 1);
                if (an < PT_MFMA_MAX_ANCHORS) {
                    const int pa = anch[gi * PT_MFMA_MAX_ANCHORS + an];   // per lane
                    const int oxa = (int)(int16_t)(pa & 0xffff), oya = pa >> 16;
                    const int xs = x0 + oxa - 8 + j, ys = y + oya - R + dyi;
                    // a key below INV_LO is a candidate inside b; the check keeps the stored
                    // match inside b whatever the key holds
                    if (xs >= 0 && ys >= 0 && xs < bew && ys < beh) {
                        mykey = (uint64_t)(uint32_t)PT_XY_TO_INT(xs, ys);
                        mys = key >> 7;   // |p|^2 + OFF - 2 q.p
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 4; t++) best[t] = ~0u;
        }
    }
    // ---- phase 3: a member whose best candidate differs from its match takes it when the
    // candidate's int8 SSD, |q|^2 + |p|^2 - 2 q.p, is strictly below the match's score (the field
    // holds int8 scores; the exact FP32 SSD of the final match is computed after the rounds).
    const int x = xw0 + lane;
    const bool live = y < ah && x < aw;
    const long vi = (long)y * aw + x;
    int v = 0;
    float d = 0.f;
    uint8_t fl = 0;
    if (live) {
        v = cur.xy[vi];
        d = cur.d[vi];
        fl = cur.fl[vi] & F_MOVED;
    }
    const int v_in = v;
    const int c = (int)(uint32_t)mykey;
    if (live && mykey != NONE && c != v) {
        const int s8 = (int)mys - (int)PT_MFMA_OFF + M.qnorm[vi];
        if ((float)s8 < d) {
            d = (float)s8;
            v = c;
            fl = F_MOVED;
        }
    }
    if (!live) return;
    nxt.xy[vi] = v;
    nxt.d[vi] = d;
    nxt.fl[vi] = fl | (v != v_in ? F_CHG : 0);
}

__global__ __launch_bounds__(MW, 5) void k_mfma_round(MfmaArgs M, NnfField cur, NnfField nxt, int J, int stable,
                                                      int seen) {
    mfma_round_body(M, cur, nxt, J, stable, seen);
}

// One round of every pair of a batch (blockIdx.z = pair), from field src into field src ^ 1.
__global__ __launch_bounds__(MW, 5) void k_mfma_round_batch(const PairLookup *__restrict__ P, int src, int J,
                                                            int stable, int seen) {
    const PairLookup &p = P[blockIdx.z];
    mfma_round_body(p.M, p.f[src], p.f[src ^ 1], J, stable, seen);
}

}  // namespace

int pt_quant_rgb0_launch(const float *img, long npx, uint32_t *out, hipStream_t s) {
    if (npx <= 0) return PT_OK;
    hipLaunchKernelGGL(k_quant_rgb0, dim3(pt_blocks(npx, 256)), dim3(256), 0, s, img, npx, out);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_mfma_prepare_launch(const float *b, int h, int w, uint32_t *bq, uint32_t *bnorm, hipStream_t s) {
    const long npx = (long)h * w, n = (long)(w - 7) * (h - 7);
    hipLaunchKernelGGL(k_quant_rgb0_pad, dim3(pt_blocks(npx, 256)), dim3(256), 0, s, b, h, w, bq);
    hipLaunchKernelGGL(k_patch_norms_pad, dim3(pt_blocks(n, 256)), dim3(256), 0, s, bq, w, w - 7, h - 7, bnorm);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

long pt_mfma_round_workgroups(int ann_w, int ann_h) { return (long)((ann_w + TILE_X - 1) / TILE_X) * ann_h; }

int pt_mfma_rounds_launch(const MfmaArgs &M, int rounds, const int *jumps, bool seen_ok, NnfField f0, NnfField f1,
                          hipStream_t s, int *which) {
    if (M.nanchors < 1 || M.nanchors > PT_MFMA_MAX_ANCHORS)
        return pt_failf(PT_E_INVALID, "mfma_anchors must be in 1..%d", PT_MFMA_MAX_ANCHORS);
    if (seen_ok && !M.anch) return pt_fail(PT_E_INVALID, "MFMA rounds: no anchor buffer");
    dim3 grid((M.ann_w + TILE_X - 1) / TILE_X, M.ann_h);
    const long nwg = pt_mfma_round_workgroups(M.ann_w, M.ann_h);
    for (int r = 0; r < rounds; r++) {
        MfmaArgs Mr = M;
        if (M.items_out) Mr.items_out = M.items_out + r * nwg;
        const int J = pt_mfma_jump_at(jumps, rounds, r), stable = r > 0 && pt_mfma_jump_at(jumps, rounds, r - 1) == J;
        hipLaunchKernelGGL(k_mfma_round, grid, dim3(MW), 0, s, Mr, f0, f1, J, stable, (int)(seen_ok && r > 0));
        std::swap(f0, f1);
    }
    PT_LAUNCH_CHECK();
    *which = rounds % 2;
    return PT_OK;
}

int pt_mfma_rounds_batch_launch(const PairLookup *P_dev, const PairLookup &P0, int npairs, int rounds, const int *jumps,
                                hipStream_t s,
                                int *which) {
    const MfmaArgs &M = P0.M;
    if (M.nanchors < 1 || M.nanchors > PT_MFMA_MAX_ANCHORS)
        return pt_failf(PT_E_INVALID, "mfma_anchors must be in 1..%d", PT_MFMA_MAX_ANCHORS);
    if (M.items_out) return pt_fail(PT_E_INVALID, "batched MFMA rounds keep no item counts");
    if (!M.anch) return pt_fail(PT_E_INVALID, "batched MFMA rounds: no anchor buffer");
    dim3 grid((M.ann_w + TILE_X - 1) / TILE_X, M.ann_h, npairs);
    int src = 0;
    for (int r = 0; r < rounds; r++) {
        const int J = pt_mfma_jump_at(jumps, rounds, r), stable = r > 0 && pt_mfma_jump_at(jumps, rounds, r - 1) == J;
        // a batch has no previous fields: every score is the int8 SSD of its match
        hipLaunchKernelGGL(k_mfma_round_batch, grid, dim3(MW), 0, s, P_dev, src, J, stable, (int)(r > 0));
        src ^= 1;
    }
    PT_LAUNCH_CHECK();
    *which = src;
    return PT_OK;
}
