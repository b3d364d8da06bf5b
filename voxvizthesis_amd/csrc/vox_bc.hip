// vox_bc.hip -- block compression of the bricks for compressed .gvb output:
// colours as BC3 (DXT5), normals as BC1 (DXT1), each (B+2)^3 brick slice by
// slice in 4x4 blocks.
//
// The reference had the OpenGL driver compress each brick as a 3D S3TC
// texture (VoxelBrickWriter.cu 590-723), so its bytes are driver-specific.
// This encoder is the project's own and deterministic (integer only):
//   endpoints  per-channel max and min of the block's texels, quantized to
//              RGB565 as (v*31+127)/255 and (v*63+127)/255; c0 >= c1;
//   palette    expanded endpoints (q<<3|q>>2, q<<2|q>>4) and the rounded
//              thirds (2a+b+1)/3, (a+2b+1)/3;
//   indices    nearest palette entry in squared RGB distance, lowest index on
//              ties; c0 == c1 gives all indices 0;
//   alpha      a0 = max, a1 = min, entries ((8-k)a0+(k-1)a1+3)/7 for k=2..7,
//              nearest, lowest on ties; a0 == a1 gives all indices 0.
// Blocks of a slice are row-major, slices follow in z (the layout of a
// compressed 3D texture image); texels past the brick edge repeat the edge.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vox_internal.h"

namespace voxmi {
namespace {

__device__ __forceinline__ uint32_t q5(uint32_t v) { return (v * 31u + 127u) / 255u; }
__device__ __forceinline__ uint32_t q6(uint32_t v) { return (v * 63u + 127u) / 255u; }
__device__ __forceinline__ uint32_t e5(uint32_t q) { return (q << 3) | (q >> 2); }
__device__ __forceinline__ uint32_t e6(uint32_t q) { return (q << 2) | (q >> 4); }

// 8-byte colour block of 16 RGB texels (r, g, b in bytes 0..2 of each word).
__device__ uint2 encode_color(const uint32_t (&t)[16]) {
    uint32_t mn[3] = {255u, 255u, 255u}, mx[3] = {0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t v = (t[i] >> (8 * c)) & 255u;
            mn[c] = v < mn[c] ? v : mn[c];
            mx[c] = v > mx[c] ? v : mx[c];
        }
    uint32_t c0 = (q5(mx[0]) << 11) | (q6(mx[1]) << 5) | q5(mx[2]);
    uint32_t c1 = (q5(mn[0]) << 11) | (q6(mn[1]) << 5) | q5(mn[2]);
    if (c0 < c1) {
        const uint32_t s = c0;
        c0 = c1;
        c1 = s;
    }
    uint32_t idx = 0u;
    if (c0 != c1) {
        int32_t P[4][3];
        P[0][0] = (int32_t)e5(c0 >> 11);
        P[0][1] = (int32_t)e6((c0 >> 5) & 63u);
        P[0][2] = (int32_t)e5(c0 & 31u);
        P[1][0] = (int32_t)e5(c1 >> 11);
        P[1][1] = (int32_t)e6((c1 >> 5) & 63u);
        P[1][2] = (int32_t)e5(c1 & 31u);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            P[2][c] = (2 * P[0][c] + P[1][c] + 1) / 3;
            P[3][c] = (P[0][c] + 2 * P[1][c] + 1) / 3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int32_t r = (int32_t)(t[i] & 255u), g = (int32_t)((t[i] >> 8) & 255u),
                          b = (int32_t)((t[i] >> 16) & 255u);
            int32_t best = 0x7fffffff;
            uint32_t bi = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int32_t dr = r - P[k][0], dg = g - P[k][1], db = b - P[k][2];
                const int32_t d = dr * dr + dg * dg + db * db;
                if (d < best) {
                    best = d;
                    bi = (uint32_t)k;
                }
            }
            idx |= bi << (2 * i);
        }
    }
    return make_uint2(c0 | (c1 << 16), idx);
}

// 8-byte alpha block of 16 alphas (byte 3 of each word).
__device__ uint2 encode_alpha(const uint32_t (&t)[16]) {
    uint32_t a0 = 0u, a1 = 255u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t a = t[i] >> 24;
        a0 = a > a0 ? a : a0;
        a1 = a < a1 ? a : a1;
    }
    uint64_t bits = 0ull;
    if (a0 != a1) {
        int32_t A[8];
        A[0] = (int32_t)a0;
        A[1] = (int32_t)a1;
#pragma unroll
        for (int k = 2; k < 8; ++k) A[k] = ((8 - k) * (int32_t)a0 + (k - 1) * (int32_t)a1 + 3) / 7;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int32_t a = (int32_t)(t[i] >> 24);
            int32_t best = 0x7fffffff;
            uint32_t bi = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int32_t d = a > A[k] ? a - A[k] : A[k] - a;
                if (d < best) {
                    best = d;
                    bi = (uint32_t)k;
                }
            }
            bits |= (uint64_t)bi << (3 * i);
        }
    }
    const uint64_t lo = (uint64_t)a0 | ((uint64_t)a1 << 8) | (bits << 16);
    return make_uint2((uint32_t)lo, (uint32_t)(lo >> 32));
}

// One thread per 4x4 block of one slice of one brick.
__global__ void k_bc_bricks(const uint32_t* __restrict__ col, const uint32_t* __restrict__ nrm,
                            uint64_t nbricks, uint32_t ex, uint32_t ey, uint32_t ez,
                            uint32_t* __restrict__ outCol, uint32_t* __restrict__ outNrm) {
    const uint32_t bxn = (ex + 3u) / 4u, byn = (ey + 3u) / 4u;
    const uint64_t perBrick = (uint64_t)bxn * byn * ez;
    const uint64_t total = nbricks * perBrick;
    const uint64_t nv = (uint64_t)ex * ey * ez;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < total;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = k / perBrick;
        uint32_t r = (uint32_t)(k - b * perBrick);
        const uint32_t z = r / (bxn * byn);
        r -= z * bxn * byn;
        const uint32_t by = r / bxn, bx = r - by * bxn;
        const uint32_t* sc = col + b * nv + (uint64_t)z * ex * ey;
        const uint32_t* sn = nrm + b * nv + (uint64_t)z * ex * ey;
        uint32_t tc[16], tn[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t x = bx * 4u + (uint32_t)(i & 3), y = by * 4u + (uint32_t)(i >> 2);
            x = x < ex ? x : ex - 1u;
            y = y < ey ? y : ey - 1u;
            tc[i] = sc[(uint64_t)y * ex + x];
            tn[i] = sn[(uint64_t)y * ex + x];
        }
        const uint2 a = encode_alpha(tc);
        const uint2 c = encode_color(tc);
        reinterpret_cast<uint4*>(outCol)[k] = make_uint4(a.x, a.y, c.x, c.y);
        reinterpret_cast<uint2*>(outNrm)[k] = encode_color(tn);
    }
}

}  // namespace

uint64_t bc_blocks_per_brick(const uint32_t e[3]) {
    return (uint64_t)((e[0] + 3u) / 4u) * ((e[1] + 3u) / 4u) * e[2];
}

void launch_bc_bricks(hipStream_t st, const uint32_t* col, const uint32_t* nrm, uint64_t nbricks,
                      const uint32_t e[3], uint32_t* outCol, uint32_t* outNrm) {
    const uint64_t total = nbricks * bc_blocks_per_brick(e);
    if (total == 0) return;
    const uint64_t g = (total + 255) / 256;
    k_bc_bricks<<<(unsigned)(g < 65536 ? g : 65536), 256, 0, st>>>(col, nrm, nbricks, e[0], e[1],
                                                                    e[2], outCol, outNrm);
}

}  // namespace voxmi
