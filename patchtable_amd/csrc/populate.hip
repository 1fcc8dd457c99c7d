// Table populate on gfx950 (patchtable.h:3121-3159, TABLE_POPULATE_FAST with populate_random).
//
// The reference lists the database patches in raster order, shuffles the list with
// std::random_shuffle driven by rand() (glibc, seed 1), and writes each patch into its cell if
// the cell is still empty. The winner of a cell is therefore the patch of minimum shuffled
// position ("rank") among the patches mapping to it. This file computes every patch's rank in
// parallel and resolves cells with atomicMin.
//
// Rank of a patch. random_shuffle performs, for i = 1..n-1, swap(a[i], a[j_i]) with
// j_i = rand_i % (i+1) (libstdc++ bits/stl_algo.h:4568-4584). Position i is untouched before
// step i, so element i lands at j_i at step i (element 0 sits at 0 from the start, j_0 = 0).
// An element at position p after step t moves at the next step m > t with j_m = p, and then sits
// at position m. Hence, with
//     G(k) = min{ m > k : j_m = j_k }     (next hit of the landing position)
//     F(q) = min{ m > q : j_m = q }       (next hit of position q once occupied at step q)
// the final position of element k is j_k if G(k) does not exist, else the end of the chain
// G(k), F(G(k)), F(F(G(k))), ... . G and F come from one stable sort of (j_m, m). F(q) > q, so
// every chain strictly increases and ends; position q is hit after step q about ln(n/q) times
// (step m hits it with probability 1/(m+1)), so chains are short and each lane walks its own.
#include "pt_device.h"
#include "pt_kernels.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace {

constexpr int RAND_CHUNK = 256;   // rand() outputs generated per thread

// j_m = rand_{m-1} % (m+1) for every shuffle step m (j_0 = 0), and m itself. Thread c jumps the
// 31-word generator state to output c*RAND_CHUNK with the matrices A^(RAND_CHUNK*2^i) and then
// steps the recurrence r[k] = r[k-31] + r[k-3] in blocks of 31 so every state index is static.
__global__ void k_shuffle_targets(const uint32_t *__restrict__ jump, const uint32_t *__restrict__ state0, int n,
                                  uint32_t *__restrict__ j, int32_t *__restrict__ m) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const long first = (long)c * RAND_CHUNK;   // output index
    const long nout = (long)n - 1;             // outputs used by steps 1..n-1
    if (c == 0) {
        j[0] = 0;
        m[0] = 0;
    }
    if (first >= nout) return;
    uint32_t st[31];
#pragma unroll
    for (int t = 0; t < 31; t++) st[t] = state0[t];
    for (int lv = 0; (c >> lv) != 0; lv++) {
        if (!((c >> lv) & 1)) continue;
        const uint32_t *P = jump + (size_t)lv * 961;
        uint32_t nx[31];
#pragma unroll
        for (int r = 0; r < 31; r++) {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < 31; k++) acc += P[r * 31 + k] * st[k];
            nx[r] = acc;
        }
#pragma unroll
        for (int t = 0; t < 31; t++) st[t] = nx[t];
    }
    const long last = first + RAND_CHUNK < nout ? first + RAND_CHUNK : nout;
    for (long base = first; base < last; base += 31) {
#pragma unroll
        for (int t = 0; t < 31; t++) {
            st[t] = st[t] + st[(t + 28) % 31];
            long o = base + t;
            if (o < last) {
                uint32_t step = (uint32_t)(o + 1);
                j[step] = (st[t] >> 1) % (step + 1);
                m[step] = (int32_t)step;
            }
        }
    }
}

__global__ void k_list_xy(int bew, int n, int32_t *__restrict__ list) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) list[k] = PT_XY_TO_INT(7 + k % bew, 7 + k / bew);
}

// From the sorted (j, m) pairs: G of every element and F of every occupied position.
__global__ void k_next_hits(const uint32_t *__restrict__ js, const int32_t *__restrict__ ms, int n,
                            int32_t *__restrict__ Gk, int32_t *__restrict__ F) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t q = js[t];
    bool has_next = t + 1 < n && js[t + 1] == q;
    Gk[ms[t]] = has_next ? ms[t + 1] : -1;
    if (t == 0 || js[t - 1] != q) {
        int32_t f = -1;
        if ((uint32_t)ms[t] > q) f = ms[t];
        else if (has_next) f = ms[t + 1];
        F[q] = f;
    }
}

// End of the chain q, F(q), F(F(q)), ...: the first position whose F is -1. F strictly
// increases along the chain, so the walk takes fewer than n steps.
__global__ void k_chain_end(const int32_t *__restrict__ F, int n, int32_t *__restrict__ fin) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int32_t c = q;
    for (int it = 0; it < n; it++) {
        const int32_t f = F[c];
        if (f < 0) break;
        c = f;
    }
    fin[q] = c;
}

__global__ void k_ranks(const int32_t *__restrict__ Gk, const int32_t *__restrict__ fin, const uint32_t *__restrict__ j,
                        int n, int32_t *__restrict__ rank) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rank[k] = Gk[k] >= 0 ? fin[Gk[k]] : (int32_t)j[k];
}

__global__ void k_identity(int n, int32_t *__restrict__ rank) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rank[k] = k;
}

// Cell index of each listed patch: sum_i get_bin_i(desc_i) * stride_i (patchtable.h:3132-3134).
__global__ void k_patch_cells(const float *__restrict__ desc, size_t plane, int Wp, const int32_t *__restrict__ list_xy,
                              int n, int G, const float *__restrict__ pmin, const float *__restrict__ prange,
                              const int32_t *__restrict__ to_bin, const int64_t *__restrict__ stride,
                              int64_t *__restrict__ cell) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int v = list_xy[k];
    size_t off = (size_t)PT_INT_TO_Y(v) * Wp + PT_INT_TO_X(v);
    int64_t idx = 0;
    for (int i = 0; i < G; i++) {
        int b = pt_get_bin(desc[i * plane + off], pmin[i], prange[i], to_bin + (size_t)i * PT_NMAP);
        idx += (int64_t)b * stride[i];
    }
    cell[k] = idx;
}

__global__ void k_claim(const int64_t *__restrict__ cell, const int32_t *__restrict__ rank, int n, int32_t *__restrict__ table) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    atomicMin(&table[cell[k]], rank[k]);
}

__global__ void k_winners(const int64_t *__restrict__ cell, const int32_t *__restrict__ rank, int n,
                          const int32_t *__restrict__ table, uint8_t *__restrict__ win) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    win[k] = table[cell[k]] == rank[k];
}

__global__ void k_write_winners(const int64_t *__restrict__ cell, const uint8_t *__restrict__ win,
                                const int32_t *__restrict__ list_xy, int n, int32_t *__restrict__ table,
                                int *__restrict__ count) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int won = 0;
    if (k < n && win[k]) {
        table[cell[k]] = list_xy[k];
        won = 1;
    }
    pt_block_count_add(won, count);
}

__global__ void k_fill(int32_t *__restrict__ p, long n, int32_t v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

int pt_list_launch(int bew, int n, int32_t *list, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_list_xy, dim3(pt_blocks(n, 256)), dim3(256), 0, s, bew, n, list);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_shuffle_ranks(const uint32_t *jump_dev, const uint32_t *state_dev, int n, int populate_random, PtScratch &sc,
                     int32_t *rank, hipStream_t s) {
    const int TB = 256;
    if (n <= 0) return PT_OK;
    if (!populate_random) {
        hipLaunchKernelGGL(k_identity, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, n, rank);
        PT_LAUNCH_CHECK();
        return PT_OK;
    }
    uint32_t *j = (uint32_t *)sc.get(0, (size_t)n * 4);
    int32_t *m = (int32_t *)sc.get(1, (size_t)n * 4);
    uint32_t *js = (uint32_t *)sc.get(2, (size_t)n * 4);
    int32_t *ms = (int32_t *)sc.get(3, (size_t)n * 4);
    int32_t *Gk = (int32_t *)sc.get(4, (size_t)n * 4);
    int32_t *F = (int32_t *)sc.get(5, (size_t)n * 4);
    int32_t *s0 = (int32_t *)sc.get(6, (size_t)n * 4);
    if (!j || !m || !js || !ms || !Gk || !F || !s0) return pt_fail(PT_E_NOMEM, "populate: scratch allocation failed");
    const long nchunks = ((long)n - 1 + RAND_CHUNK - 1) / RAND_CHUNK;
    hipLaunchKernelGGL(k_shuffle_targets, dim3(pt_blocks(nchunks > 0 ? nchunks : 1, 64)), dim3(64), 0, s, jump_dev,
                       state_dev, n, j, m);
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)n) bits++;
    size_t tmp_bytes = 0;
    PT_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, j, js, m, ms, n, 0, bits, s));
    void *tmp = sc.get(7, tmp_bytes);
    if (!tmp) return pt_fail(PT_E_NOMEM, "populate: sort scratch allocation failed");
    PT_HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, j, js, m, ms, n, 0, bits, s));
    PT_HIP_TRY(hipMemsetAsync(F, 0xff, (size_t)n * 4, s));
    hipLaunchKernelGGL(k_next_hits, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, js, ms, n, Gk, F);
    hipLaunchKernelGGL(k_chain_end, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, F, n, s0);
    hipLaunchKernelGGL(k_ranks, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, Gk, s0, j, n, rank);
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_populate_launch(const float *desc, size_t plane, int Wp, const int32_t *list_xy, const int32_t *rank, int n,
                       const DevQuant &q, int32_t *table, int64_t ncells, int64_t *cell_scratch,
                       uint8_t *win_scratch, int *count_dev, hipStream_t s) {
    const int TB = 256;
    hipLaunchKernelGGL(k_fill, dim3(pt_blocks(ncells, TB)), dim3(TB), 0, s, table, (long)ncells, (int32_t)PT_TABLE_UNUSED);
    PT_HIP_TRY(hipMemsetAsync(count_dev, 0, sizeof(int), s));
    if (n > 0) {
        hipLaunchKernelGGL(k_patch_cells, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, desc, plane, Wp, list_xy, n, q.G,
                           q.min_val, q.range, q.to_bin, q.stride, cell_scratch);
        hipLaunchKernelGGL(k_claim, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, cell_scratch, rank, n, table);
        hipLaunchKernelGGL(k_winners, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, cell_scratch, rank, n, table, win_scratch);
        hipLaunchKernelGGL(k_write_winners, dim3(pt_blocks(n, TB)), dim3(TB), 0, s, cell_scratch, win_scratch, list_xy,
                           n, table, count_dev);
    }
    PT_LAUNCH_CHECK();
    return PT_OK;
}

int pt_fill_launch(int32_t *p, int64_t n, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(pt_blocks(n, 256)), dim3(256), 0, s, p, (long)n, v);
    PT_LAUNCH_CHECK();
    return PT_OK;
}
