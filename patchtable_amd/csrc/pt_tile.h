// Tiled patch-SSD helpers shared by the lookup and k-coherence kernels (gfx950).
//
// A workgroup of TILE_X x TILE_Y lanes owns a tile of 8x8 patches (top-left corners) and stages
// the tile's (TILE_X+7) x (TILE_Y+7) pixels in LDS, one float4 (R, G, B, 0) per pixel; the other
// patch of each SSD is read from global memory with unaligned 16-byte loads.
#ifndef PT_TILE_H
#define PT_TILE_H

#include "pt_device.h"

namespace pt_tile {

constexpr int TILE_X = 64, TILE_Y = 4, THREADS = TILE_X * TILE_Y;
constexpr int AX = TILE_X + 7, AY = TILE_Y + 7;   // staged query pixels

// Tile (column, row) of this workgroup. Workgroups go to the 8 XCDs round-robin in linear
// dispatch order (x fastest), so plain blockIdx gives each XCD every 8th tile column. When the
// columns divide evenly, XCD k instead takes the band of gridDim.x / 8 adjacent columns from
// column k * gridDim.x / 8, row by row: neighbouring tiles, which share halo pixels and
// candidate lines, then share one L2. The map is a bijection of the grid.
__device__ __forceinline__ int2 tile_coords() {
    const int gx = gridDim.x;
    if (gx % 8 != 0) return make_int2(blockIdx.x, blockIdx.y);
    const int lin = blockIdx.y * gx + blockIdx.x, band = gx / 8, j = lin / 8;
    return make_int2((lin % 8) * band + j % band, j / band);
}

// Stages the pixels of this block's tile of patches into LDS, zero outside the image. The tile
// is TX x (THREADS / TX) patches, (TX + 7) x (THREADS / TX + 7) pixels.
template <int TX = TILE_X>
__device__ __forceinline__ void stage_tile(const float *__restrict__ img, int ih, int iw, float4 *tile) {
    constexpr int TY = THREADS / TX, SX = TX + 7, SY = TY + 7;
    const int2 tc = tile_coords();
    const int x0 = tc.x * TX, y0 = tc.y * TY;
    for (int i = threadIdx.x; i < SX * SY; i += THREADS) {
        int ly = i / SX, lx = i - ly * SX;
        int gx = x0 + lx, gy = y0 + ly;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gx < iw && gy < ih) {
            const float *p = img + ((size_t)gy * iw + gx) * 3;
            v = make_float4(p[0], p[1], p[2], 0.f);
        }
        tile[i] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ float4 ldg4(const float *p) {
    float4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}
__device__ __forceinline__ int4 ldg4i(const int32_t *p) {
    int4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

#define PT_ACC(e, bv)                                  \
    {                                                  \
        float d_ = av[e] - (bv);                       \
        acc[(e) & 3] = acc[(e) & 3] + d_ * d_;         \
    }

// Exact SSD of the staged query patch at `aq` (row pitch PITCH pixels) against the database patch
// at `b` (row pitch `bpitch` floats), lane order of patch_dist_exact (patchtable.h:330-373):
// element e of a 24-float row goes to lane e & 3, lanes accumulate in increasing e then row.
// With `bounded`, the sum is abandoned once a partial sum (after any row but the last) reaches
// `bound`.
// Rows are processed one at a time (83 VGPRs, 5 waves per SIMD).
template <bool bounded, int PITCH = AX>
__device__ __forceinline__ float ssd_staged(const float4 *aq, const float *__restrict__ b, int bpitch, float bound) {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int dy = 0; dy < 8; dy++) {
        const float4 *ar = aq + dy * PITCH;
        const float *br = b + (size_t)dy * bpitch;
        float4 b0 = ldg4(br), b1 = ldg4(br + 4), b2 = ldg4(br + 8), b3 = ldg4(br + 12), b4 = ldg4(br + 16),
               b5 = ldg4(br + 20);
        float av[24];
#pragma unroll
        for (int px = 0; px < 8; px++) {
            float4 p = ar[px];
            av[3 * px + 0] = p.x;
            av[3 * px + 1] = p.y;
            av[3 * px + 2] = p.z;
        }
        PT_ACC(0, b0.x) PT_ACC(1, b0.y) PT_ACC(2, b0.z) PT_ACC(3, b0.w)
        PT_ACC(4, b1.x) PT_ACC(5, b1.y) PT_ACC(6, b1.z) PT_ACC(7, b1.w)
        PT_ACC(8, b2.x) PT_ACC(9, b2.y) PT_ACC(10, b2.z) PT_ACC(11, b2.w)
        PT_ACC(12, b3.x) PT_ACC(13, b3.y) PT_ACC(14, b3.z) PT_ACC(15, b3.w)
        PT_ACC(16, b4.x) PT_ACC(17, b4.y) PT_ACC(18, b4.z) PT_ACC(19, b4.w)
        PT_ACC(20, b5.x) PT_ACC(21, b5.y) PT_ACC(22, b5.z) PT_ACC(23, b5.w)
        if (bounded && dy < 7) {
            float part = ((acc[0] + acc[1]) + acc[2]) + acc[3];
            if (part >= bound) return part;
        }
    }
    return ((acc[0] + acc[1]) + acc[2]) + acc[3];
}
#undef PT_ACC

}  // namespace pt_tile

#endif
