// k-coherence lists on gfx950 (replaces the FLANN kd-tree search of patchtable.h:2653-2881).
//
// For every database patch p the reference keeps its k nearest database patches in descriptor
// space that lie at pixel distance >= kcoherence_min_dist, re-scored by exact RGB SSD and sorted
// by it. FLANN's arithmetic is third-party and unpinned; this library finds candidates with a
// data-parallel approximate kNN instead:
//   1. every patch gets two 64-bit keys, the Morton interleave of its table bins and of its
//      bins offset by half the slice count; patches are radix-sorted by each key (ties by
//      raster index);
//   2. the candidates of p are the patches within `window` sorted positions of p in either
//      order; those at pixel distance^2 < min_dist^2 are dropped;
//   3. the K candidates nearest in descriptor L2 (all dims, summed in index order; ties to the
//      smaller raster index, duplicates removed) are kept in a register top-K list;
//   4. they are re-scored by the lookup distance (exact RGB SSD; patch_dist_approx of the
//      descriptors in descriptor mode) and sorted by (distance, XY_TO_INT) as the reference sorts
//      its (dist, (index, valid)) triples (patchtable.h:2815-2835).
// Lists shorter than K repeat their last entry (p itself with distance 0 when empty). With an
// allowed mask only allowed patches are listed and searched (TABLE_FOR_ALLOWED_PATCHES_STEP,
// patchtable.h:2711); a patch outside the mask gets K entries (p, +inf), which end a walk at once.
#include "pt_device.h"
#include "pt_kernels.h"
#include "pt_tile.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace {

struct KeyArgs {
    int G;          // grid dims
    int Gk;         // dims interleaved in the key
    int B;          // bits per dim
    int half[PT_MAX_DIMS];
};

// Keys carry a flag above the Morton bits for patches outside the mask, which sorts them last.
// Each patch's nd descriptor values are also written side by side (dI), for the sorted gathers.
__global__ void k_keys(const float *__restrict__ desc, size_t plane, int Wp, int bew, long N, DevQuant q, KeyArgs ka,
                       const int32_t *__restrict__ allowed, int aidx, uint64_t *__restrict__ k0,
                       uint64_t *__restrict__ k1, int32_t *__restrict__ idx, int nd, float *__restrict__ dI) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int px = (int)(i % bew), py = (int)(i / bew);
    size_t off = (size_t)(py + 7) * Wp + px + 7;
    for (int c = 0; c < nd; c++) dI[i * nd + c] = desc[c * plane + off];
    uint32_t bins[PT_MAX_DIMS];
    for (int d = 0; d < ka.Gk; d++)
        bins[d] = (uint32_t)pt_get_bin(desc[d * plane + off], q.min_val[d], q.range[d], q.to_bin + (size_t)d * PT_NMAP);
    uint64_t a = 0, b = 0;
    for (int bit = ka.B - 1; bit >= 0; bit--)
        for (int d = 0; d < ka.Gk; d++) {
            a = (a << 1) | ((bins[d] >> bit) & 1u);
            b = (b << 1) | (((bins[d] + (uint32_t)ka.half[d]) >> bit) & 1u);
        }
    if (allowed && allowed[i] != aidx) {
        a |= 1ull << (ka.Gk * ka.B);
        b |= 1ull << (ka.Gk * ka.B);
    }
    k0[i] = a;
    k1[i] = b;
    idx[i] = (int32_t)i;
}

__global__ void k_count_allowed(const int32_t *__restrict__ allowed, int aidx, long N, int *__restrict__ cnt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    pt_block_count_add(i < N && allowed[i] == aidx, cnt);
}

__global__ void k_positions(const int32_t *__restrict__ perm, long N, int32_t *__restrict__ pos) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) pos[perm[r]] = (int32_t)r;
}

// Descriptors in both sorted orders: ds0[r] = dI[perm0[r]], ds1[r] = dI[perm1[r]] (nd floats each).
template <int ND>
__global__ void k_gather_desc(const float *__restrict__ dI, int nd_rt, const int32_t *__restrict__ perm0,
                              const int32_t *__restrict__ perm1, long N, float *__restrict__ ds0,
                              float *__restrict__ ds1) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int nd = ND > 0 ? ND : nd_rt;
    const float *a = dI + (size_t)perm0[r] * nd, *b = dI + (size_t)perm1[r] * nd;
    float *oa = ds0 + (size_t)r * nd, *ob = ds1 + (size_t)r * nd;
    if constexpr (ND == 6) {   // 24-byte rows: three 8-byte loads each
#pragma unroll
        for (int c = 0; c < 6; c += 2) {
            const float2 va = *(const float2 *)(a + c), vb = *(const float2 *)(b + c);
            *(float2 *)(oa + c) = va;
            *(float2 *)(ob + c) = vb;
        }
    } else {
        for (int c = 0; c < nd; c++) {
            oa[c] = a[c];
            ob[c] = b[c];
        }
    }
}

// A (distance, index) pair as one 64-bit key that orders by distance, then by index. The
// distances here are sums of squares (+0.0 or positive, never -0.0), whose bit patterns order as
// the floats do; the index's sign bit is flipped so that signed indices keep their order. With a
// single integer comparison per step the top-K insertions below are plain min/max swaps: the
// equivalent compound test (d < d' || (d == d' && i < i')) was miscompiled for gfx950 in the
// selection loop (on a distance tie the slot kept its old entry while the carry took it too,
// leaving a duplicate in the list).
__device__ __forceinline__ uint64_t dist_key(float d, int32_t i) {
    return ((uint64_t)__float_as_uint(d) << 32) | ((uint32_t)i ^ 0x80000000u);
}
__device__ __forceinline__ int32_t key_index(uint64_t k) { return (int32_t)((uint32_t)k ^ 0x80000000u); }
__device__ __forceinline__ float key_dist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }

// One lane per database patch p, in 64 x 4 raster tiles whose pixels are staged in LDS for the
// exact-SSD re-rank (DESC: re-rank by descriptor distance, nothing staged). ND is the descriptor
// width when known at compile time (0: runtime `nd`). Na: allowed patches (sorted first).
template <int KMAX, int ND, bool DESC>
__global__ __launch_bounds__(pt_tile::THREADS) void k_select(
    const int32_t *__restrict__ perm0, const int32_t *__restrict__ perm1, const int32_t *__restrict__ pos0,
    const int32_t *__restrict__ pos1, const float *__restrict__ ds0, const float *__restrict__ ds1, long Na, int bew,
    int beh, int nd_rt, int K, int W, int md2, const float *__restrict__ b0, const float *__restrict__ bI, int bw,
    int bh, const int32_t *__restrict__ allowed, int aidx, int32_t *__restrict__ kpos, float *__restrict__ kdist) {
    using namespace pt_tile;
    __shared__ float4 tile[DESC ? 1 : AX * AY];
    if (!DESC) stage_tile(b0, bh, bw, tile);
    const int lx = threadIdx.x % TILE_X, ly = threadIdx.x / TILE_X;
    const int2 tc = tile_coords();
    const int px = tc.x * TILE_X + lx, py = tc.y * TILE_Y + ly;
    if (px >= bew || py >= beh) return;
    const int nd = ND > 0 ? ND : nd_rt;
    const int i = py * bew + px;
    if (allowed && allowed[i] != aidx) {
        for (int k = 0; k < K; k++) {
            kpos[(size_t)i * K + k] = PT_XY_TO_INT(px, py);
            kdist[(size_t)i * K + k] = INFINITY;
        }
        return;
    }
    float own[ND > 0 ? ND : PT_MAX_DIMS];
    const long r0 = pos0[i];
    for (int c = 0; c < nd; c++) own[c] = ds0[r0 * nd + c];
    uint64_t bk[KMAX];   // dist_key(descriptor distance, raster index), ascending; ~0 when empty
#pragma unroll
    for (int k = 0; k < KMAX; k++) bk[k] = ~0ull;
    int cnt = 0;
    uint64_t wk = ~0ull;   // K-th entry once the list is full
    for (int o = 0; o < 2; o++) {
        const int32_t *perm = o == 0 ? perm0 : perm1;
        const float *ds = o == 0 ? ds0 : ds1;
        const long r = o == 0 ? r0 : (long)pos1[i];
        const long t0 = r - W < 0 ? 0 : r - W, t1 = r + W >= Na ? Na - 1 : r + W;
        for (long t4 = t0; t4 <= t1; t4 += 4) {
          // the window is contiguous in the sorted arrays: with ND known, four candidates' ranks
          // and descriptors arrive in 16-byte loads (one for the ranks, ND for the values)
          int qv[4];
          float dv[4][ND > 0 ? ND : 1];
          const bool full = ND > 0 && t4 + 3 <= t1;
          if (full) {
              const int4 p4 = pt_tile::ldg4i(perm + t4);
              qv[0] = p4.x;
              qv[1] = p4.y;
              qv[2] = p4.z;
              qv[3] = p4.w;
              if constexpr (ND > 0) {
                  float tmp[4 * ND];
#pragma unroll
                  for (int j = 0; j < ND; j++) {
                      const float4 v = pt_tile::ldg4(ds + t4 * ND + 4 * j);
                      tmp[4 * j] = v.x;
                      tmp[4 * j + 1] = v.y;
                      tmp[4 * j + 2] = v.z;
                      tmp[4 * j + 3] = v.w;
                  }
#pragma unroll
                  for (int u = 0; u < 4; u++)
#pragma unroll
                      for (int c = 0; c < ND; c++) dv[u][c] = tmp[u * ND + c];
              }
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const long t = t4 + u;
            if (t > t1) break;
            if (t == r) continue;
            const int qi = full ? qv[u] : perm[t];
            const int qx = qi % bew, qy = qi / bew;
            const int dx = qx - px, dy = qy - py;
            if (dx * dx + dy * dy < md2) continue;
            float dd = 0.f;
            if (full) {
                if constexpr (ND > 0) {
#pragma unroll
                    for (int c = 0; c < ND; c++) {
                        float e = own[c] - dv[u][c];
                        dd = dd + e * e;
                    }
                }
            } else {
                const float *dq = ds + t * nd;
                for (int c = 0; c < nd; c++) {
                    float e = own[c] - dq[c];
                    dd = dd + e * e;
                }
            }
            // reject if not better than the current K-th entry
            const uint64_t key = dist_key(dd, qi);
            if (cnt == K && !(key < wk)) continue;
            bool dup = false;
#pragma unroll
            for (int k = 0; k < KMAX; k++) dup |= (k < cnt && key_index(bk[k]) == qi);
            if (dup) continue;
            // insert keeping the keys ascending (the empty slots hold ~0, above every key); a
            // full list drops its last key
            const int n = cnt < K ? cnt + 1 : K;
            uint64_t ck = key;
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
                if (k < n) {
                    const uint64_t t = bk[k];
                    bk[k] = ck < t ? ck : t;
                    ck = ck < t ? t : ck;
                }
            }
            cnt = n;
#pragma unroll
            for (int k = 0; k < KMAX; k++)
                if (k == K - 1) wk = bk[k];
          }
        }
    }
    // exact re-rank, then insertion by (distance, XY_TO_INT)
    const float4 *aq = tile + (DESC ? 0 : ly * AX + lx);
    const int Wpb = bw + 7;
    const float *own_i = bI ? bI + ((size_t)(py + 7) * Wpb + px + 7) * nd : nullptr;
    uint64_t sk[KMAX];   // dist_key(lookup distance, XY_TO_INT), ascending; ~0 when empty
#pragma unroll
    for (int k = 0; k < KMAX; k++) sk[k] = ~0ull;
#pragma unroll
    for (int a = 0; a < KMAX; a++) {
        if (a < cnt) {
            const int q = key_index(bk[a]), qx = q % bew, qy = q / bew;
            const float cs = DESC ? pt_desc_dist(own_i, bI + ((size_t)(qy + 7) * Wpb + qx + 7) * nd, nd)
                                  : ssd_staged<false>(aq, b0 + ((size_t)qy * bw + qx) * 3, bw * 3, 0.f);
            uint64_t ck = dist_key(cs, PT_XY_TO_INT(qx, qy));
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
                if (k <= a) {
                    const uint64_t t = sk[k];
                    sk[k] = ck < t ? ck : t;
                    ck = ck < t ? t : ck;
                }
            }
        }
    }
    int32_t *op = kpos + (size_t)i * K;
    float *od = kdist + (size_t)i * K;
    int lastxy = PT_XY_TO_INT(px, py);
    float lasts = 0.f;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < cnt) {
            lastxy = key_index(sk[k]);
            lasts = key_dist(sk[k]);
        }
        if (k < K) {
            op[k] = lastxy;
            od[k] = lasts;
        }
    }
}

}  // namespace

int pt_kcoherence_launch(const float *desc, size_t plane, int Wp, int ndims, const float *b0, const float *b_descI,
                         int bw, int bh, const int32_t *allowed, int aidx, const DevQuant &q, const int *nslices_host,
                         int K, int window, int min_dist2, int32_t *kpos, float *kdist, PtScratch &sc, hipStream_t s) {
    const int bew = bw - 7, beh = bh - 7;
    const long N = (long)bew * beh;
    if (K <= 0 || N <= 0) return PT_OK;
    if (K > 32) return pt_failf(PT_E_UNSUPPORTED, "kcoherence %d > 32 is not supported", K);
    if (ndims > PT_MAX_DIMS) return pt_fail(PT_E_UNSUPPORTED, "too many dims");
    KeyArgs ka;
    ka.G = q.G;
    int mx = 1;
    for (int d = 0; d < q.G; d++) {
        int v = nslices_host[d] + (nslices_host[d] >> 1);
        mx = v > mx ? v : mx;
    }
    ka.B = 0;
    while ((1 << ka.B) < mx) ka.B++;
    if (ka.B == 0) ka.B = 1;
    ka.Gk = q.G;
    while (ka.Gk * ka.B > 63) ka.Gk--;   // one bit above the key for the allowed flag
    for (int d = 0; d < PT_MAX_DIMS; d++) ka.half[d] = d < q.G ? (nslices_host[d] >> 1) : 0;
    const int nbits = ka.Gk * ka.B + 1;

    uint64_t *k0 = (uint64_t *)sc.get(20, N * 8), *k1 = (uint64_t *)sc.get(21, N * 8);
    uint64_t *k0s = (uint64_t *)sc.get(22, N * 8), *k1s = (uint64_t *)sc.get(23, N * 8);
    int32_t *idx = (int32_t *)sc.get(24, N * 4);
    int32_t *perm0 = (int32_t *)sc.get(25, N * 4), *perm1 = (int32_t *)sc.get(26, N * 4);
    int32_t *pos1 = (int32_t *)sc.get(27, N * 4), *pos0 = (int32_t *)sc.get(31, N * 4);
    float *ds0 = (float *)sc.get(28, (size_t)N * ndims * 4), *ds1 = (float *)sc.get(29, (size_t)N * ndims * 4);
    if (!k0 || !k1 || !k0s || !k1s || !idx || !perm0 || !perm1 || !pos0 || !pos1 || !ds0 || !ds1)
        return pt_fail(PT_E_NOMEM, "kcoherence: scratch allocation failed");
    const int TB = 256;
    float *dI = (float *)sc.get(33, (size_t)N * ndims * 4);
    if (!dI) return pt_fail(PT_E_NOMEM, "kcoherence: scratch allocation failed");
    hipLaunchKernelGGL(k_keys, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, desc, plane, Wp, bew, N, q, ka, allowed, aidx, k0,
                       k1, idx, ndims, dI);
    long Na = N;   // allowed patches
    if (allowed) {
        int *cnt = (int *)sc.get(32, sizeof(int));
        if (!cnt) return pt_fail(PT_E_NOMEM, "kcoherence: scratch allocation failed");
        PT_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_count_allowed, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, allowed, aidx, N, cnt);
        int hc = 0;
        PT_HIP_TRY(hipMemcpyAsync(&hc, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
        PT_HIP_TRY(hipStreamSynchronize(s));
        Na = hc;
    }
    size_t tmp_bytes = 0;
    PT_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, k0, k0s, idx, perm0, (int)N, 0, nbits, s));
    void *tmp = sc.get(30, tmp_bytes);
    if (!tmp) return pt_fail(PT_E_NOMEM, "kcoherence: sort scratch allocation failed");
    PT_HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k0s, idx, perm0, (int)N, 0, nbits, s));
    PT_HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, k1, k1s, idx, perm1, (int)N, 0, nbits, s));
    hipLaunchKernelGGL(k_positions, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, perm0, N, pos0);
    hipLaunchKernelGGL(k_positions, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, perm1, N, pos1);
    if (ndims == 6)
        hipLaunchKernelGGL(k_gather_desc<6>, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, dI, ndims, perm0, perm1, N, ds0,
                           ds1);
    else
        hipLaunchKernelGGL(k_gather_desc<0>, dim3(pt_blocks(N, TB)), dim3(TB), 0, s, dI, ndims, perm0, perm1, N, ds0,
                           ds1);
    dim3 grid((bew + pt_tile::TILE_X - 1) / pt_tile::TILE_X, (beh + pt_tile::TILE_Y - 1) / pt_tile::TILE_Y);
    dim3 blk(pt_tile::THREADS);
#define PT_SELECT(KM, ND, DESC)                                                                                 \
    hipLaunchKernelGGL((k_select<KM, ND, DESC>), grid, blk, 0, s, perm0, perm1, pos0, pos1, ds0, ds1, Na, bew, beh, \
                       ndims, K, window, min_dist2, b0, b_descI, bw, bh, allowed, aidx, kpos, kdist)
    if (b_descI) {
        if (K <= 16) PT_SELECT(16, 0, true);
        else PT_SELECT(32, 0, true);
    } else if (K <= 16) {
        if (ndims == 6) PT_SELECT(16, 6, false);
        else PT_SELECT(16, 0, false);
    } else {
        if (ndims == 6) PT_SELECT(32, 6, false);
        else PT_SELECT(32, 0, false);
    }
#undef PT_SELECT
    PT_LAUNCH_CHECK();
    return PT_OK;
}
