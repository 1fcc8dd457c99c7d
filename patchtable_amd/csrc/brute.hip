// Exhaustive nearest-neighbour search over every database patch, on gfx950.
//
// The reference names LOOKUP_ALGO_BRUTE (patchtable.h:230, "-lookup_algo brute",
// patchtable.cpp:276) but has no search behind it; this file is the dense form of candidate
// scoring: for query patch a and database patch b (192-dim RGB vectors, centred by -0.5),
//     |a - b|^2 = |a|^2 + |b|^2 - 2 a.b,
// so ranking all of b for a block of queries is a [queries x 192] x [192 x database] contraction.
//
//   k_patch_vectors   patch vectors in BF16 (round to nearest even) or FP8 e4m3 with a per-patch
//                     power-of-two scale (an E8M0 block scale of the instruction), and |v|^2 of the
//                     rounded vector in FP32.
//   k_brute_wide      one block = 256 queries (8 waves x 32); the database streams through two
//                     LDS buffers in chunks of 128 patches (the next chunk's global loads overlap
//                     the current chunk's MFMAs); 32 x 32 output tiles: BF16 a.b on
//                     v_mfma_f32_32x32x16_bf16, or FP8 |b|^2 - 2 a.b directly on
//                     v_mfma_scale_f32_32x32x64_f8f6f4 (K = 192 in three steps of 64; accumulators
//                     start at |b|^2, the scales are applied by the instruction), FP32 accumulate;
//                     each lane keeps the 2 best scores |b|^2 - 2 a.b of its 16 query rows over the
//                     database columns it sees; at the end the 64 lane candidates of a query are
//                     merged to its 8 best.
//   k_rescore         exact FP32 SSD (the reference's lane order, patch_dist_exact,
//                     patchtable.h:330-373) of the 8 survivors; the smallest is the match, so no
//                     low-precision error reaches the reported distance.
//   k_brute_f64 / k_f64_reduce
//                     validation only: the exact nearest neighbour in FP64 (sum over the 192
//                     elements in row-major order with fused multiply-adds), ties to the smaller
//                     database index.
#include "pt_device.h"
#include "pt_kernels.h"

#include <algorithm>

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int DIM = 192;
constexpr int QPB = 256;        // queries per block
constexpr int KL = 2;           // candidates a lane keeps per query row

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// Patch p: positions pos[p] = XY_TO_INT(x, y) (patch coordinates), or raster order over an
// ew-wide patch grid when pos is NULL.
__device__ __forceinline__ void patch_xy(const int32_t *pos, long p, int ew, int &x, int &y) {
    if (pos) {
        x = PT_INT_TO_X(pos[p]);
        y = PT_INT_TO_Y(pos[p]);
    } else {
        x = (int)(p % ew);
        y = (int)(p / ew);
    }
}

// One wave per patch; lane l holds elements l, l + 64, l + 128 (k = dy*24 + dx*3 + c).
// fp8 = 0: vec is BF16 bits (2 bytes per element). fp8 = 1 (queries) or 2 (database): vec is e4m3
// bytes with a power-of-two scale[p] = 2^e, the smallest with max |v| / scale <= 448, so the
// matrix cores can apply it as an E8M0 block scale; database vectors (fp8 = 2) hold -v / scale
// and scale[p] = 2 * 2^e, so the contraction yields -2 a.b. norm[p] is |v|^2 of the rounded vector.
__global__ __launch_bounds__(256) void k_patch_vectors(const float *__restrict__ img, int w, const int32_t *__restrict__ pos,
                                                      long n, int ew, int fp8, unsigned char *__restrict__ vec,
                                                      float *__restrict__ norm, float *__restrict__ scale) {
    const long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= n) return;
    int x, y;
    patch_xy(pos, p, ew, x, y);
    float v[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int k = lane + 64 * i;
        v[i] = img[((size_t)(y + k / 24) * w + x) * 3 + k % 24] - 0.5f;
    }
    float s = 0.f;
    if (!fp8) {
        unsigned short *o = (unsigned short *)vec + (size_t)p * DIM;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const unsigned short b = f2bf(v[i]);
            o[lane + 64 * i] = b;
            const float r = bf2f(b);
            s += r * r;
        }
    } else {
        float m = fmaxf(fabsf(v[0]), fmaxf(fabsf(v[1]), fabsf(v[2])));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        // m / 448 = f 2^e with f in [0.5, 1): scale 2^e keeps every |v| / scale <= 448
        int e = 0;
        (void)frexpf(m > 0.f ? m / 448.f : 1.f, &e);
        e = min(max(e, -126), 127);
        const float sc = ldexpf(1.f, e);
        const float inv = ldexpf(1.f, -e);
        const float sgn = fp8 == 2 ? -1.f : 1.f;
        unsigned char *o = vec + (size_t)p * DIM;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            // v_cvt_pk_fp8_f32 converts two values; the first of the pair is the one kept
            const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(sgn * (v[i] * inv), 0.f, 0, false);
            o[lane + 64 * i] = (unsigned char)(packed & 0xff);
            const float r = __builtin_amdgcn_cvt_f32_fp8(packed, 0) * sc;
            s += r * r;
        }
        if (lane == 0) scale[p] = fp8 == 2 ? 2.f * sc : sc;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) norm[p] = s;
}

// Keeps the KL smallest (score, index) pairs; equal scores keep the earlier index.
__device__ __forceinline__ void keep_best(float s, int idx, float (&bs)[KL], int (&bi)[KL]) {
    if (s < bs[1]) {
        if (s < bs[0]) {
            bs[1] = bs[0];
            bi[1] = bi[0];
            bs[0] = s;
            bi[0] = idx;
        } else {
            bs[1] = s;
            bi[1] = idx;
        }
    }
}

// The 8 best of a lane's query: its candidates (mscore/midx row `row`, n entries, -1 = none) merged
// into (ts, ti), ascending, ties to the smaller index.
__device__ __forceinline__ void merge_best8(const float *mrow, const int *irow, int n, float (&ts)[8], int (&ti)[8]) {
    for (int e = 0; e < n; e++) {
        float s = mrow[e];
        int id = irow[e];
        if (id < 0) continue;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const bool better = s < ts[i] || (s == ts[i] && (ti[i] < 0 || id < ti[i]));
            if (better) {
                const float fs = ts[i];
                const int fi = ti[i];
                ts[i] = s;
                ti[i] = id;
                s = fs;
                id = fi;
                if (id < 0) break;
            }
        }
    }
}

// One block = 256 queries (8 waves x 32); the database streams through two LDS buffers in chunks
// of 128 patches (4 column tiles of 32; the next chunk's global loads overlap the current chunk's
// MFMAs). Lane l holds A row / B column l % 32 and, of each K step, the second half of the step's
// elements when l >= 32 (the same elements of A and B, so the products pair up); accumulator r
// holds query row 8 (r / 4) + 4 (l / 32) + r % 4 against database column l % 32.
//   BF16 (FP8 = false): a.b on v_mfma_f32_32x32x16_bf16, twelve K steps of 16, FP32 accumulate;
//     score |b|^2 - 2 a.b.
//   FP8: |b|^2 - 2 a.b straight from v_mfma_scale_f32_32x32x64_f8f6f4: K = 192 is three steps of
//     64 with no padding, the accumulators start at |b|^2 and the E8M0 block scales (one per row
//     of A, one per column of B) are applied by the instruction.
// Each lane keeps the KL best scores of its 16 query rows over the columns it sees; a query's
// 32 lanes x KL candidates are merged to its 8 best.
constexpr int WW = 8;            // waves per block
constexpr int WDBC = 128;        // database patches per chunk
template <bool FP8>
__global__ __launch_bounds__(64 * WW, 2) void k_brute_wide(const unsigned char *__restrict__ qvec,
                                                           const float *__restrict__ qscale, int nq,
                                                           const unsigned char *__restrict__ dvec,
                                                           const float *__restrict__ dnorm,
                                                           const float *__restrict__ dscale, int nd,
                                                           int32_t *__restrict__ cand, float *__restrict__ cscore) {
    constexpr int NT = 64 * WW;
    constexpr int EB = FP8 ? 1 : 2;            // bytes per element
    constexpr int NKS = FP8 ? 3 : 12;          // K steps
    constexpr int SB = DIM * EB / NKS;         // bytes of one K step (64 or 32), half per lane group
    constexpr int ROWB = DIM * EB + 16;        // LDS row pitch: 16 bytes against bank conflicts
    constexpr int MWW = WW / 2;                // waves whose lists are merged at a time
    constexpr int MERGE = MWW * 32 * 32 * KL;  // their lists: 32 rows x 32 lanes x KL each
    constexpr int SMEM = 2 * WDBC * ROWB > MERGE * 8 ? 2 * WDBC * ROWB : MERGE * 8;
    static_assert(QPB == 32 * WW, "one row tile of 32 queries per wave");
    typedef int frag_t __attribute__((ext_vector_type(SB / 8)));   // a lane's half step
    __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];
    __shared__ float snorm[2][WDBC];
    __shared__ float sscale[2][FP8 ? WDBC : 1];
    float (*mscore)[32][32 * KL] = (float (*)[32][32 * KL])smem;
    int (*midx)[32][32 * KL] = (int (*)[32][32 * KL])(smem + MERGE * 4);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l32 = lane & 31, kh = lane >> 5;
    const int qw0 = blockIdx.x * QPB + wave * 32;   // first query of this wave
    auto half_step = [&](const unsigned char *row, int ks) {
        frag_t f;
        __builtin_memcpy(&f, row + SB * ks + (SB / 2) * kh, SB / 2);
        return f;
    };
    // A fragments of query qw0 + l32 (clamped to the last query; such rows are never written)
    // and, for FP8, the E8M0 exponent of its scale
    const int qa = min(qw0 + l32, nq - 1);
    frag_t afr[NKS];
#pragma unroll
    for (int ks = 0; ks < NKS; ks++) afr[ks] = half_step(qvec + (size_t)qa * DIM * EB, ks);
    const int qe = FP8 ? (int)((__float_as_uint(qscale[qa]) >> 23) & 0xffu) : 0;
    float bs[16][KL];
    int bi[16][KL];
#pragma unroll
    for (int r = 0; r < 16; r++)
#pragma unroll
        for (int m = 0; m < KL; m++) {
            bs[r][m] = __builtin_inff();
            bi[r][m] = -1;
        }
    // staging: thread i moves 16-byte pieces i, i + NT, ... of a chunk (zero rows past the end)
    constexpr int PIECES = WDBC * DIM * EB / 16, PER = PIECES / NT, RP = DIM * EB / 16;
    static_assert(PIECES % NT == 0, "chunk pieces");
    v4i stage[PER];
    float stn = 0.f, sts = 1.f;
    auto fetch = [&](int d0) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + NT * k, row = i / RP, c16 = i % RP, d = d0 + row;
            stage[k] = d < nd ? *(const v4i *)(dvec + (size_t)d * DIM * EB + c16 * 16) : v4i{0, 0, 0, 0};
        }
        if (threadIdx.x < WDBC) {
            const int d = d0 + threadIdx.x;
            stn = d < nd ? dnorm[d] : __builtin_inff();
            if (FP8) sts = d < nd ? dscale[d] : 1.f;
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = threadIdx.x + NT * k, row = i / RP, c16 = i % RP;
            *(v4i *)(smem + (size_t)buf * WDBC * ROWB + row * ROWB + c16 * 16) = stage[k];
        }
        if (threadIdx.x < WDBC) {
            snorm[buf][threadIdx.x] = stn;
            if (FP8) sscale[buf][threadIdx.x] = sts;
        }
    };
    fetch(0);
    put(0);
    __syncthreads();
    for (int d0 = 0, cur = 0; d0 < nd; d0 += WDBC, cur ^= 1) {
        const bool more = d0 + WDBC < nd;
        if (more) fetch(d0 + WDBC);   // loads in flight during this chunk's MFMAs
        const unsigned char *sdb = smem + (size_t)cur * WDBC * ROWB;
#pragma unroll 1
        for (int c = 0; c < WDBC / 32; c++) {
            const int dcol = c * 32 + l32;
            const int didx = d0 + dcol;
            const float dn = snorm[cur][dcol];
            const unsigned char *brow = sdb + dcol * ROWB;
            v16f acc;
            if constexpr (FP8) {
                const int be = (int)((__float_as_uint(sscale[cur][dcol]) >> 23) & 0xffu);
#pragma unroll
                for (int r = 0; r < 16; r++) acc[r] = dn;
#pragma unroll
                for (int ks = 0; ks < NKS; ks++)   // e4m3 operands (cbsz = blgp = 0), scale byte 0 of each
                    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(afr[ks], half_step(brow, ks), acc, 0, 0, 0,
                                                                          qe, 0, be);
            } else {
#pragma unroll
                for (int r = 0; r < 16; r++) acc[r] = 0.f;
#pragma unroll
                for (int ks = 0; ks < NKS; ks++)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, afr[ks]),
                                                                 __builtin_bit_cast(bf8, half_step(brow, ks)), acc, 0, 0,
                                                                 0);
#pragma unroll
                for (int r = 0; r < 16; r++) acc[r] = dn - 2.f * acc[r];
            }
            // a score enters a list only when it beats the list's second entry; once the lists
            // have seen a few thousand columns that is rare, so the wave tests all 16 first and
            // runs the insertions only when some lane needs one
            bool enter = false;
#pragma unroll
            for (int r = 0; r < 16; r++) enter = enter || acc[r] < bs[r][KL - 1];
            if (__any(enter)) {
#pragma unroll
                for (int r = 0; r < 16; r++) keep_best(acc[r], didx, bs[r], bi[r]);
            }
        }
        if (more) put(cur ^ 1);   // the other buffer was last read before the previous barrier
        __syncthreads();
    }
    // merge, MWW waves at a time: the lane lists of query row 8 (r / 4) + 4 kh + r % 4 go to
    // mscore[wave % MWW][row][l32 * KL + m]; then lane l < 32 takes query qw0 + l
    const int q = qw0 + l32;
    float ts[8];
    int ti[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        ts[i] = __builtin_inff();
        ti[i] = -1;
    }
#pragma unroll 1
    for (int phase = 0; phase < WW / MWW; phase++) {
        __syncthreads();   // the database chunk, or the previous phase's lists, are no longer read
        const bool mine = wave / MWW == phase;
        const int mw = wave % MWW;
        if (mine) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int m = 0; m < KL; m++) {
                    const int row = 8 * (r >> 2) + 4 * kh + (r & 3);
                    mscore[mw][row][l32 * KL + m] = bs[r][m];
                    midx[mw][row][l32 * KL + m] = bi[r][m];
                }
        }
        __syncthreads();
        if (mine && kh == 0 && q < nq) merge_best8(mscore[mw][l32], midx[mw][l32], 32 * KL, ts, ti);
    }
    if (kh != 0 || q >= nq) return;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        cand[(size_t)q * 8 + i] = ti[i];
        cscore[(size_t)q * 8 + i] = ts[i];
    }
}

// Exact SSD of the 8 survivors of each query; the smallest (ties: better approximate rank) is
// the match. out_pos = XY_TO_INT(x, y) in patch coordinates.
__global__ void k_rescore(const float *__restrict__ a, int aw, const int32_t *__restrict__ qpos, int nq, int qew,
                          const float *__restrict__ b, int bw, int dew, const int32_t *__restrict__ cand,
                          int32_t *__restrict__ out_pos, float *__restrict__ out_dist) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int qx, qy;
    patch_xy(qpos, q, qew, qx, qy);
    const float *ap = a + ((size_t)qy * aw + qx) * 3;
    float best = __builtin_inff();
    int bpos = -1;
    for (int i = 0; i < 8; i++) {
        const int d = cand[(size_t)q * 8 + i];
        if (d < 0) continue;
        const int dx = d % dew, dy = d / dew;
        const float s = pt_patch_ssd(ap, 3 * aw, b + ((size_t)dy * bw + dx) * 3, 3 * bw);
        if (s < best) {
            best = s;
            bpos = PT_XY_TO_INT(dx, dy);
        }
    }
    out_pos[q] = bpos;
    out_dist[q] = best;
}

// ---- FP64 exact nearest neighbour (validation) -------------------------------------------------
constexpr int F64Q = 16;   // queries per block

__global__ __launch_bounds__(256) void k_brute_f64(const float *__restrict__ a, int aw, const int32_t *__restrict__ qpos,
                                                  int nq, int qew, const float *__restrict__ b, int bw, int dew,
                                                  long nd, int per_block, double *__restrict__ pdist,
                                                  int32_t *__restrict__ pidx) {
    __shared__ double sq[F64Q][DIM];
    __shared__ double rd[F64Q][4];
    __shared__ int ri[F64Q][4];
    const int q0 = blockIdx.y * F64Q;
    for (int i = threadIdx.x; i < F64Q * DIM; i += blockDim.x) {
        const int qi = i / DIM, k = i % DIM;
        const int q = min(q0 + qi, nq - 1);
        int qx, qy;
        patch_xy(qpos, q, qew, qx, qy);
        sq[qi][k] = (double)a[((size_t)(qy + k / 24) * aw + qx) * 3 + k % 24];
    }
    __syncthreads();
    double best[F64Q];
    int bidx[F64Q];
#pragma unroll
    for (int qi = 0; qi < F64Q; qi++) {
        best[qi] = 1e300;
        bidx[qi] = -1;
    }
    const long dbeg = (long)blockIdx.x * per_block, dend = min(nd, dbeg + per_block);
    for (long d = dbeg + threadIdx.x; d < dend; d += blockDim.x) {
        const int dx = (int)(d % dew), dy = (int)(d / dew);
        double acc[F64Q];
#pragma unroll
        for (int qi = 0; qi < F64Q; qi++) acc[qi] = 0.0;
        for (int r = 0; r < 8; r++) {
            const float *br = b + ((size_t)(dy + r) * bw + dx) * 3;
            double bv[24];
#pragma unroll
            for (int e = 0; e < 24; e++) bv[e] = (double)br[e];
#pragma unroll
            for (int qi = 0; qi < F64Q; qi++)
#pragma unroll
                for (int e = 0; e < 24; e++) {
                    const double df = sq[qi][r * 24 + e] - bv[e];
                    acc[qi] = __builtin_fma(df, df, acc[qi]);
                }
        }
#pragma unroll
        for (int qi = 0; qi < F64Q; qi++)
            if (acc[qi] < best[qi]) {   // d increases along the loop: ties keep the smaller index
                best[qi] = acc[qi];
                bidx[qi] = (int)d;
            }
    }
    // minimum over the wave (ties: smaller index), then over the 4 waves
#pragma unroll
    for (int qi = 0; qi < F64Q; qi++) {
        double v = best[qi];
        int ix = bidx[qi];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off);
            const int oi = __shfl_xor(ix, off);
            if (oi >= 0 && (ix < 0 || ov < v || (ov == v && oi < ix))) {
                v = ov;
                ix = oi;
            }
        }
        if ((threadIdx.x & 63) == 0) {
            rd[qi][threadIdx.x >> 6] = v;
            ri[qi][threadIdx.x >> 6] = ix;
        }
    }
    __syncthreads();
    if (threadIdx.x < F64Q) {
        const int qi = threadIdx.x;
        double bd = 1e300;
        int bx = -1;
        for (int t = 0; t < 4; t++) {
            const double v = rd[qi][t];
            const int ix = ri[qi][t];
            if (ix >= 0 && (v < bd || (v == bd && ix < bx))) {
                bd = v;
                bx = ix;
            }
        }
        const int q = q0 + qi;
        if (q < nq) {
            pdist[(size_t)blockIdx.x * nq + q] = bd;
            pidx[(size_t)blockIdx.x * nq + q] = bx;
        }
    }
}

__global__ void k_f64_reduce(const double *__restrict__ pdist, const int32_t *__restrict__ pidx, int nblk, int nq,
                             int dew, int32_t *__restrict__ out_pos, double *__restrict__ out_dist) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double bd = 1e300;
    int bx = -1;
    for (int k = 0; k < nblk; k++) {
        const double v = pdist[(size_t)k * nq + q];
        const int ix = pidx[(size_t)k * nq + q];
        if (ix >= 0 && (v < bd || (v == bd && ix < bx))) {
            bd = v;
            bx = ix;
        }
    }
    out_pos[q] = bx < 0 ? -1 : PT_XY_TO_INT(bx % dew, bx / dew);
    out_dist[q] = bd;
}

}  // namespace

int pt_brute_vectors_launch(const float *img, int w, const int32_t *pos, long n, int ew, int fp8, unsigned char *vec,
                            float *norm, float *scale, hipStream_t s) {
    if (n <= 0) return PT_OK;
    hipLaunchKernelGGL(k_patch_vectors, dim3(pt_blocks(n, 4)), dim3(256), 0, s, img, w, pos, n, ew, fp8, vec, norm,
                       scale);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_brute_mfma_launch(const unsigned char *qvec, const float *qscale, int nq, const unsigned char *dvec,
                         const float *dnorm, const float *dscale, int nd, int fp8, int32_t *cand, float *cscore,
                         hipStream_t s) {
    if (nq <= 0) return PT_OK;
    if (fp8)
        hipLaunchKernelGGL(k_brute_wide<true>, dim3(pt_blocks(nq, QPB)), dim3(64 * WW), 0, s, qvec, qscale, nq, dvec,
                           dnorm, dscale, nd, cand, cscore);
    else
        hipLaunchKernelGGL(k_brute_wide<false>, dim3(pt_blocks(nq, QPB)), dim3(64 * WW), 0, s, qvec, qscale, nq, dvec,
                           dnorm, dscale, nd, cand, cscore);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_brute_rescore_launch(const float *a, int aw, const int32_t *qpos, int nq, int qew, const float *b, int bw,
                            int dew, const int32_t *cand, int32_t *out_pos, float *out_dist, hipStream_t s) {
    if (nq <= 0) return PT_OK;
    hipLaunchKernelGGL(k_rescore, dim3(pt_blocks(nq, 256)), dim3(256), 0, s, a, aw, qpos, nq, qew, b, bw, dew, cand,
                       out_pos, out_dist);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_brute_f64_blocks(long nd, int *per_block) {
    const int nblk = (int)std::min<long>(64, (nd + 255) / 256);
    *per_block = (int)((nd + nblk - 1) / nblk);
    return nblk;
}

int pt_brute_f64_launch(const float *a, int aw, const int32_t *qpos, int nq, int qew, const float *b, int bw, int dew,
                        long nd, double *pdist, int32_t *pidx, int32_t *out_pos, double *out_dist, hipStream_t s) {
    if (nq <= 0) return PT_OK;
    int per = 0;
    const int nblk = pt_brute_f64_blocks(nd, &per);
    hipLaunchKernelGGL(k_brute_f64, dim3(nblk, (nq + F64Q - 1) / F64Q), dim3(256), 0, s, a, aw, qpos, nq, qew, b, bw,
                       dew, nd, per, pdist, pidx);
    hipLaunchKernelGGL(k_f64_reduce, dim3(pt_blocks(nq, 256)), dim3(256), 0, s, pdist, pidx, nblk, nq, dew, out_pos,
                       out_dist);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
