// vox_kernels.hip -- HIP kernels of the MI355X voxelizer (gfx950).
//
// Pipeline (one stream, everything resident in HBM):
//   setup      per triangle: face normal, voxel bounds, count of 8^3 tiles
//   scatter    (group, tile) bins of triangle indices (counting sort)
//   activity   tiles with triangles, plus the terrain-fill reach below them
//   voxelize   one workgroup per active tile and group: the tile's triangles
//              are staged in LDS with their plane / edge-test constants, the
//              (triangle, voxel) pairs of their clipped bounds are spread over
//              the workgroup, hits are reduced per voxel with LDS atomics
//              (count, three lowest and the highest triangle index), then each
//              hit voxel averages colour and normal
//   fill       terrain fill below textured hits
//   mip        2x2x2 box filter, tile to tile
//   node_type  per octree node: brick + 1-voxel border equal to its sample?
//   assign / gather / collapse: brick addresses by prefix sum, brick payload,
//              constant-branch collapse
//
// All fp32 arithmetic follows the reference's operation order (GLM 0.9.3
// dot / cross / normalize) and is compiled with -ffp-contract=off; division
// and sqrt are correctly rounded (hipcc default).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "vox_internal.h"

#pragma clang fp contract(off)

namespace voxmi {
namespace {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

__device__ __forceinline__ float cget(const f3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 muls3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 x, f3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
__device__ __forceinline__ f3 normalize3(f3 v) {
    float sqr = v.x * v.x + v.y * v.y + v.z * v.z;
    float inv = 1.0f / sqrtf(sqr);
    return muls3(v, inv);
}
__device__ __forceinline__ float gmin(float x, float y) { return x < y ? x : y; }
__device__ __forceinline__ float gmax(float x, float y) { return x > y ? x : y; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmax(gmin(x, hi), lo); }

__device__ __forceinline__ int64_t trunc_i(float f) {
    if (!(f == f)) return 0;
    if (f >= 9.0e18f) return INT64_C(9000000000000000000);
    if (f <= -9.0e18f) return -INT64_C(9000000000000000000);
    return (int64_t)f;
}
__device__ __forceinline__ uint32_t quant_u8(float f) {
    if (!(f == f)) return 0u;
    if (f <= 0.0f) return 0u;
    if (f >= 255.0f) return 255u;
    return (uint32_t)(int)f;
}
__device__ __forceinline__ uint32_t pack4(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return r | (g << 8) | (b << 16) | (a << 24);
}
// VoxColorToVec4 (VoxelizerKernels.cu 657-663): c / 255.0f per channel,
// read from a 256-entry table of the correctly rounded quotients (filled by
// fill_inv255 at kernel start; a multiply by 1/255 differs in 126 entries).
__device__ __forceinline__ f4 unpack4(uint32_t c, const float* __restrict__ inv255) {
    return {inv255[c & 255u], inv255[(c >> 8) & 255u], inv255[(c >> 16) & 255u],
            inv255[c >> 24]};
}
__device__ __forceinline__ void fill_inv255(float* s_inv255) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_inv255[i] = (float)i / 255.0f;
}

// Linear tile index -> tile coordinates (32-bit arithmetic; tile grids hold
// fewer than 2^31 tiles).
__device__ __forceinline__ void tile_xyz(uint64_t tlin64, const uint32_t* TG, uint32_t& x,
                                         uint32_t& y, uint32_t& z) {
    const uint32_t t = (uint32_t)tlin64;
    const uint32_t q = t / TG[0];
    x = t - q * TG[0];
    z = q / TG[1];
    y = q - z * TG[1];
}
// Node index at a depth with Rd = 2^ld nodes per axis.
__device__ __forceinline__ void node_xyz(uint64_t n, uint32_t ld, uint32_t& x, uint32_t& y,
                                         uint32_t& z) {
    const uint64_t m = (1ull << ld) - 1ull;
    x = (uint32_t)(n & m);
    y = (uint32_t)((n >> ld) & m);
    z = (uint32_t)(n >> (2 * ld));
}
__device__ __forceinline__ uint32_t to_color(f4 c) {
    return pack4(quant_u8(c.x * 255.0f), quant_u8(c.y * 255.0f), quant_u8(c.z * 255.0f),
                 quant_u8(c.w * 255.0f));
}

__device__ __forceinline__ f3 load3(const float* p) { return {p[0], p[1], p[2]}; }

// Brick normal encoding of WriteUByteBinaryBrick (VoxelBrickWriter.cu
// 557-568): ((n + 1) * 0.5) * 255 truncated, RGB8 in the low three bytes.
__device__ __forceinline__ uint32_t quant_normal(f3 n) {
    float qx = ((n.x + 1.0f) * 0.5f), qy = ((n.y + 1.0f) * 0.5f), qz = ((n.z + 1.0f) * 0.5f);
    return pack4(quant_u8(qx * 255.0f), quant_u8(qy * 255.0f), quant_u8(qz * 255.0f), 0u);
}
constexpr uint32_t kZeroNormalQ = 0x007f7f7fu;   // quant_normal of (0, 0, 0)

// Inclusive scan across a 64-wide wavefront.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        uint32_t s = s_w[i];
        pre += (i < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Walk over the (group, tile) bins of a box of tiles, x fastest.
struct TileWalk {
    uint32_t x0, x1, y0, y1, z1, x, y, z;
    uint32_t TGx, TGy;
    uint64_t base, key;
    __device__ TileWalk(const VoxParams& P, int g, uint32_t ax0, uint32_t ax1, uint32_t ay0,
                        uint32_t ay1, uint32_t az0, uint32_t az1)
        : x0(ax0), x1(ax1), y0(ay0), y1(ay1), z1(az1), x(ax0), y(ay0), z(az0), TGx(P.TG[0]),
          TGy(P.TG[1]), base((uint64_t)g * P.numTiles) {
        key = base + ((uint64_t)z * TGy + y) * TGx + x;
    }
    // advance; false past the last bin
    __device__ bool next() {
        if (++x > x1) {
            x = x0;
            if (++y > y1) {
                y = y0;
                if (++z > z1) return false;
            }
        }
        key = base + ((uint64_t)z * TGy + y) * TGx + x;
        return true;
    }
};

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int l) {
    const uint32_t lo = __shfl((uint32_t)v, l), hi = __shfl((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// The active lanes of a wave grouped by key: the group's first lane
// (leader), a lane's rank in its group (lanes in ascending order) and, in
// the leader, the group's size. Ballots and shuffles only; one pass per
// distinct key.
struct BinGroup {
    int leader;
    uint32_t rank, count;
    bool lead;
};
__device__ __forceinline__ BinGroup bin_group(bool act, uint64_t key) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint64_t below = (1ull << lane) - 1ull;
    BinGroup r{lane, 0u, 0u, false};
    uint64_t pend = __ballot(act);
    while (pend) {
        const int l = __ffsll((unsigned long long)pend) - 1;
        const uint64_t lk = shfl_u64(key, l);
        const bool in = act && key == lk;
        const uint64_t same = __ballot(in);
        if (in) {
            r.leader = l;
            r.rank = (uint32_t)__popcll(same & below);
        }
        if (lane == l) {
            r.count = (uint32_t)__popcll(same);
            r.lead = true;
        }
        pend &= ~same;
    }
    return r;
}

// i / d for i < 2^22 and a small divisor d, given inv = 1.0f / d: the float
// quotient is within one of the integer quotient; one correction step.
__device__ __forceinline__ uint32_t udiv_small(uint32_t i, uint32_t d, float inv) {
    uint32_t q = (uint32_t)((float)i * inv);
    const int32_t r = (int32_t)(i - q * d);
    if (r < 0) --q;
    else if (r >= (int32_t)d) ++q;
    return q;
}

// Binning works on (triangle, tile) pairs: a workgroup lists the pairs of
// its triangles with a block scan of their tile-box sizes, then every lane
// takes one pair, so a wave's lanes stay busy whatever the spread of box
// sizes. A triangle whose box has more than kExpandMax tiles walks its box
// on its own lane instead (the pair count of a workgroup stays < 2^22).
constexpr uint32_t kExpandMax = 4096;
struct PairList {
    uint32_t off[kBlock + 1];   // exclusive scan of pairs per triangle; off[kBlock] = total
    uint32_t xy[kBlock];        // tile box start x | y << 16
    uint32_t z[kBlock];         // tile box start z
    uint32_t ext[kBlock];       // tile box extent x | y << 16
    uint32_t tri[kBlock];       // triangle index
    uint32_t w[kBlock / 64];
};
// The bin key of pair p < off[kBlock], and its triangle.
__device__ __forceinline__ uint64_t pair_key(const PairList& L, const VoxParams& P, uint64_t base,
                                             uint32_t p, uint32_t& tri) {
    uint32_t lo = 0, hi = kBlock;   // last triangle j with off[j] <= p (its box holds p)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L.off[mid] <= p) lo = mid;
        else hi = mid;
    }
    const uint32_t r = p - L.off[lo], e = L.ext[lo];
    const uint32_t ex = e & 0xffffu, exy = ex * (e >> 16);
    const uint32_t dz = udiv_small(r, exy, 1.0f / (float)exy);
    const uint32_t rxy = r - dz * exy;
    const uint32_t dy = udiv_small(rxy, ex, 1.0f / (float)ex);
    const uint32_t dx = rxy - dy * ex, xy = L.xy[lo];
    tri = L.tri[lo];
    return base + ((uint64_t)(L.z[lo] + dz) * P.TG[1] + ((xy >> 16) + dy)) * P.TG[0] +
           (xy & 0xffffu) + dx;
}
// Voxel bounds (packed min | max << 16 per axis) clipped to the group's
// reach (its z range; x and y the chunks it overlaps): false when empty.
__device__ __forceinline__ bool clip_box(const GroupArgs& G, uint32_t bx, uint32_t by,
                                         uint32_t bz, uint32_t (&v0)[3], uint32_t (&v1)[3]) {
    v0[0] = max(bx & 0xffffu, G.xStart); v1[0] = min(bx >> 16, G.xEnd);
    v0[1] = max(by & 0xffffu, G.yStart); v1[1] = min(by >> 16, G.yEnd);
    v0[2] = max(bz & 0xffffu, G.zStart); v1[2] = min(bz >> 16, G.zEnd);
    return v0[0] <= v1[0] && v0[1] <= v1[1] && v0[2] <= v1[2];
}
// Tile box of those clipped bounds: false when empty.
__device__ __forceinline__ bool tile_box(const GroupArgs& G, uint32_t bx, uint32_t by,
                                         uint32_t bz, uint32_t (&t0)[3], uint32_t (&t1)[3]) {
    uint32_t v0[3], v1[3];
    if (!clip_box(G, bx, by, bz, v0, v1)) return false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        t0[a] = v0[a] >> 3;
        t1[a] = v1[a] >> 3;
    }
    return true;
}
// Fill the workgroup's pair list (barriers inside: every thread calls it).
// Returns the number of pairs; sets big for a box walked by its own lane.
__device__ __forceinline__ uint32_t list_pairs(PairList& L, bool has, const uint32_t (&t0)[3],
                                               const uint32_t (&t1)[3], uint32_t tri, bool& big) {
    uint32_t w = 0;
    big = false;
    if (has) {
        const uint64_t n = (uint64_t)(t1[0] - t0[0] + 1) * (t1[1] - t0[1] + 1) * (t1[2] - t0[2] + 1);
        if (n <= kExpandMax) w = (uint32_t)n;
        else big = true;
    }
    uint32_t total;
    const uint32_t off = block_excl_scan(w, L.w, &total);   // barriers
    const int tid = threadIdx.x;
    L.off[tid] = off;
    L.xy[tid] = t0[0] | (t0[1] << 16);
    L.z[tid] = t0[2];
    L.ext[tid] = w ? ((t1[0] - t0[0] + 1) | ((t1[1] - t0[1] + 1) << 16)) : 1u | (1u << 16);
    L.tri[tid] = tri;
    if (tid == 0) L.off[kBlock] = total;
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------
// setup: ComputeEdgesFaceNormalsAndBounds (VoxelizerKernels.cu 230-347) plus
// the count of 8^3 tiles each triangle's bounds touch, per (group, tile) bin.
// Bounds are packed as min | max << 16 per axis.
// The workgroup's group of a launch over several groups (scalar registers).
__device__ __forceinline__ int set_group(const GroupSet& S) {
    int gi = 0;
    while (gi + 1 < S.n && blockIdx.x >= S.start[gi + 1]) ++gi;
    return __builtin_amdgcn_readfirstlane(gi);
}

__global__ __launch_bounds__(kBlock) void k_setup(VoxParams P, GroupSet S,
                                                  const float* __restrict__ verts,
                                                  float* __restrict__ fnorm,
                                                  uint32_t* __restrict__ bbox,
                                                  uint32_t* __restrict__ binCount) {
    __shared__ PairList L;
    const int gi = set_group(S);
    const GroupArgs& G = S.G[gi];
    const int g = S.g[gi];
    const uint64_t lt = (uint64_t)(blockIdx.x - S.start[gi]) * blockDim.x + threadIdx.x;
    const uint64_t t = G.triOff + lt;
    uint32_t t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
    bool has = false;
    if (lt < G.ntri) {
        const float* q = verts + 9 * t;
        f3 v[3] = {load3(q), load3(q + 3), load3(q + 6)};
        f3 e0 = sub3(v[1], v[0]);
        f3 p0p2 = sub3(v[2], v[0]);
        f3 n = normalize3(cross3(e0, p0p2));
        fnorm[3 * t] = n.x;
        fnorm[3 * t + 1] = n.y;
        fnorm[3 * t + 2] = n.z;
        uint32_t mn[3], mx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f3 diff = {(v[k].x - P.offP[0]) / P.dP[0], (v[k].y - P.offP[1]) / P.dP[1],
                       (v[k].z - P.offP[2]) / P.dP[2]};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int64_t iv = trunc_i(cget(diff, a));
                int64_t hi = (int64_t)P.S[a] - 1;
                uint32_t u = (uint32_t)(iv < 0 ? 0 : (iv > hi ? hi : iv));
                if (k == 0) { mn[a] = u; mx[a] = u; }
                else { mn[a] = u < mn[a] ? u : mn[a]; mx[a] = u > mx[a] ? u : mx[a]; }
            }
        }
        uint32_t pk[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            pk[a] = mn[a] | (mx[a] << 16);
            bbox[3 * t + a] = pk[a];
        }
        has = G.overlaps && tile_box(G, pk[0], pk[1], pk[2], t0, t1);
    }
    if (!G.overlaps) return;   // uniform: the whole workgroup
    bool big;
    const uint32_t total = list_pairs(L, has, t0, t1, (uint32_t)t, big);
    // Neighbouring triangles mostly share tiles: the lanes holding the same
    // bin add once, with the count of the group, all groups of a step in one
    // atomic instruction.
    const uint64_t base = (uint64_t)g * P.numTiles;
    for (uint32_t p0 = 0; p0 < total; p0 += kBlock) {
        const uint32_t p = p0 + threadIdx.x;
        const bool act = p < total;
        uint32_t tri = 0;
        const uint64_t key = act ? pair_key(L, P, base, p, tri) : 0ull;
        const BinGroup bg = bin_group(act, key);
        if (act && bg.lead) atomicAdd(&binCount[key], bg.count);
    }
    if (!__syncthreads_or(big)) return;
    TileWalk w(P, g, t0[0], t1[0], t0[1], t1[1], t0[2], t1[2]);
    bool act = big;
    while (__ballot(act)) {
        const BinGroup bg = bin_group(act, w.key);
        if (act && bg.lead) atomicAdd(&binCount[w.key], bg.count);
        if (act) act = w.next();
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(VoxParams P, GroupSet S,
                                                    const uint32_t* __restrict__ bbox,
                                                    const uint32_t* __restrict__ binStart,
                                                    uint32_t* __restrict__ binFill,
                                                    uint32_t* __restrict__ binTri) {
    const int gi = set_group(S);
    const GroupArgs& G = S.G[gi];
    const int g = S.g[gi];
    if (!G.overlaps) return;   // uniform: the whole workgroup
    __shared__ PairList L;
    const uint64_t lt = (uint64_t)(blockIdx.x - S.start[gi]) * blockDim.x + threadIdx.x;
    const uint64_t t = G.triOff + lt;
    uint32_t t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
    bool has = false;
    if (lt < G.ntri) has = tile_box(G, bbox[3 * t], bbox[3 * t + 1], bbox[3 * t + 2], t0, t1);
    bool big;
    const uint32_t total = list_pairs(L, has, t0, t1, (uint32_t)t, big);
    // One returning atomic instruction per step reserves the places of every
    // group of lanes holding the same bin; the lanes of a group take
    // consecutive places after its base.
    const uint64_t base = (uint64_t)g * P.numTiles;
    for (uint32_t p0 = 0; p0 < total; p0 += kBlock) {
        const uint32_t p = p0 + threadIdx.x;
        const bool act = p < total;
        uint32_t tri = 0;
        const uint64_t key = act ? pair_key(L, P, base, p, tri) : 0ull;
        const BinGroup bg = bin_group(act, key);
        uint32_t r = 0;
        if (act && bg.lead) r = atomicAdd(&binFill[key], bg.count);
        const uint32_t b0 = (uint32_t)__shfl((int)r, bg.leader);
        if (act) binTri[binStart[key] + b0 + bg.rank] = tri;
    }
    if (!__syncthreads_or(big)) return;
    TileWalk w(P, g, t0[0], t1[0], t0[1], t1[1], t0[2], t1[2]);
    bool act = big;
    while (__ballot(act)) {
        const BinGroup bg = bin_group(act, w.key);
        uint32_t r = 0;
        if (act && bg.lead) r = atomicAdd(&binFill[w.key], bg.count);
        const uint32_t b0 = (uint32_t)__shfl((int)r, bg.leader);
        if (act) binTri[binStart[w.key] + b0 + bg.rank] = (uint32_t)t;
        if (act) act = w.next();
    }
}

// A level-0 tile is active when any group has triangles binned to it, or when
// a terrain group with a texture has triangles up to three tiles above it
// (the fill reaches 19 voxels down, VoxelizerKernels.cu 888-917).
__global__ void k_activity0(VoxParams P, int ngroups, const int32_t* __restrict__ grpFill,
                            const uint32_t* __restrict__ binStart, uint32_t* __restrict__ mark) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.numTiles) return;
    uint32_t tx_, ty_, tz;
    tile_xyz(t, P.TG, tx_, ty_, tz);
    uint32_t on = 0;
    for (int g = 0; g < ngroups && !on; ++g) {
        uint64_t key = (uint64_t)g * P.numTiles + t;
        if (binStart[key + 1] > binStart[key]) on = 1;
        if (!on && grpFill[g]) {
            for (int q = 1; q <= kFillTiles; ++q) {
                if (tz + q >= P.TG[2]) break;
                uint64_t k2 = key + (uint64_t)q * P.TG[0] * P.TG[1];
                if (binStart[k2 + 1] > binStart[k2]) { on = 1; break; }
            }
        }
    }
    mark[t] = on;
}

__global__ void k_compact(uint64_t n, const uint32_t* __restrict__ mark,
                          const uint32_t* __restrict__ excl, int32_t* __restrict__ tileIdx,
                          int32_t* __restrict__ list, uint32_t* __restrict__ count) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *count = excl[n];   // the scan's total: active tiles
    if (t >= n) return;
    if (mark[t]) {
        uint32_t e = excl[t];
        tileIdx[t] = (int32_t)e;
        list[e] = (int32_t)t;
    } else {
        tileIdx[t] = -1;
    }
}

// Level-k tile (a,b,c) reads source voxels [16a, 16a+16), i.e. source tiles
// 2a and 2a+1 on each axis, shifted by the tile's chunk offset (MipShift,
// whole tile pairs) with several chunks.
__global__ void k_mark_children(LevelMap src, LevelMap dst, MipShift shift,
                                uint32_t* __restrict__ markDst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dst.numTiles) return;
    uint32_t px, py, pz;
    tile_xyz(t, dst.TG, px, py, pz);
    const uint32_t ox = shift.cd[0] ? (px * kTile) / shift.cd[0] * shift.off[0] / kTile : 0u;
    const uint32_t oy = shift.cd[1] ? (py * kTile) / shift.cd[1] * shift.off[1] / kTile : 0u;
    const uint32_t oz = shift.cd[2] ? (pz * kTile) / shift.cd[2] * shift.off[2] / kTile : 0u;
    uint32_t m = 0u;
#pragma unroll
    for (uint32_t c = 0; c < 8u; ++c) {
        const uint32_t x = 2 * px + ox + (c & 1u), y = 2 * py + oy + ((c >> 1) & 1u),
                       z = 2 * pz + oz + (c >> 2);
        if (x < src.TG[0] && y < src.TG[1] && z < src.TG[2] &&
            src.tileIdx[((uint64_t)z * src.TG[1] + y) * src.TG[0] + x] >= 0)
            m = 1u;
    }
    markDst[t] = m;
}

// ---------------------------------------------------------------------------
// voxelize: ComputeVoxelization (VoxelizerKernels.cu 349-530) for one group
// on one tile, then ComputeColorsAndNormals / ComputeColorsAsNormals
// (VoxelizerKernels.cu 689-1083) for the tile's hit voxels.
constexpr int kStage = 128;     // triangles staged per round
constexpr int kTriStride = 35;  // floats per staged triangle (padded, odd)

// Bilinear tex2D with normalized coordinates (VoxelizerKernels.cu 586-600):
// fractional weights held in 8 fractional bits; NaN coordinates read as 0.
__device__ __forceinline__ int wrap_index(int i, int n, int mode, bool& border) {
    if (mode == 0) { int m = i % n; return m < 0 ? m + n : m; }
    if (mode == 2) {
        int period = 2 * n;
        int m = i % period;
        if (m < 0) m += period;
        return m < n ? m : period - 1 - m;
    }
    if (mode == 3) {
        if (i < 0 || i >= n) { border = true; return 0; }
        return i;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
__device__ __forceinline__ f4 texel(const GroupArgs& G, int i, int j,
                                    const float* __restrict__ inv255) {
    bool border = false;
    int x = wrap_index(i, G.w, G.addr0, border);
    int y = wrap_index(j, G.h, G.addr1, border);
    if (border) return {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t c = *reinterpret_cast<const uint32_t*>(G.img + 4 * ((size_t)y * (size_t)G.w + (size_t)x));
    return unpack4(c, inv255);
}
__device__ f4 tex2d(const GroupArgs& G, float u, float v, const float* __restrict__ inv255) {
    if (!(u == u)) u = 0.0f;
    if (!(v == v)) v = 0.0f;
    float xb = u * (float)G.w - 0.5f;
    float yb = v * (float)G.h - 0.5f;
    float fx = floorf(xb), fy = floorf(yb);
    float ax = rintf((xb - fx) * 256.0f) / 256.0f;
    float ay = rintf((yb - fy) * 256.0f) / 256.0f;
    int i = (int)fmaxf(fminf(fx, 1.0e9f), -1.0e9f);
    int j = (int)fmaxf(fminf(fy, 1.0e9f), -1.0e9f);
    f4 t00 = texel(G, i, j, inv255), t10 = texel(G, i + 1, j, inv255);
    f4 t01 = texel(G, i, j + 1, inv255), t11 = texel(G, i + 1, j + 1, inv255);
    float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay);
    float w01 = (1.0f - ax) * ay, w11 = ax * ay;
    f4 r;
    r.x = ((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x;
    r.y = ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y;
    r.z = ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z;
    r.w = ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w;
    return r;
}

// ComputeVoxelBarycentricCoords (VoxelizerKernels.cu 602-624).
__device__ __forceinline__ f2 bary(f2 P, f2 A, f2 B, f2 C) {
    f2 v0 = {C.x - A.x, C.y - A.y};
    f2 v1 = {B.x - A.x, B.y - A.y};
    f2 v2 = {P.x - A.x, P.y - A.y};
    float dot00 = v0.x * v0.x + v0.y * v0.y;
    float dot01 = v0.x * v1.x + v0.y * v1.y;
    float dot02 = v0.x * v2.x + v0.y * v2.y;
    float dot11 = v1.x * v1.x + v1.y * v1.y;
    float dot12 = v1.x * v2.x + v1.y * v2.y;
    float invDenom = 1.0f / (dot00 * dot11 - dot01 * dot01);
    return {gclamp((dot11 * dot02 - dot01 * dot12) * invDenom, 0.0f, 1.0f),
            gclamp((dot00 * dot12 - dot01 * dot02) * invDenom, 0.0f, 1.0f)};
}
__device__ __forceinline__ f3 interp_n(f2 uv, f3 a, f3 b, f3 c) {
    f3 V = muls3(sub3(b, a), uv.y);
    f3 U = muls3(sub3(c, a), uv.x);
    return add3(add3(a, U), V);
}
__device__ __forceinline__ f2 bary_axis(int axis, f3 c, const f3* v) {
    if (axis == 0) return bary({c.z, c.y}, {v[0].z, v[0].y}, {v[1].z, v[1].y}, {v[2].z, v[2].y});
    if (axis == 1) return bary({c.x, c.z}, {v[0].x, v[0].z}, {v[1].x, v[1].z}, {v[2].x, v[2].z});
    return bary({c.x, c.y}, {v[0].x, v[0].y}, {v[1].x, v[1].y}, {v[2].x, v[2].y});
}

// The triangle attributes the colour pass reads, 27 floats: vertices [0, 9),
// face normal [9, 12), texture coordinates [12, 18) (textured groups),
// vertex normals [18, 27) (groups that have them).
constexpr int kTriAttr = 27;
__device__ __forceinline__ void load_tri_attr(const GroupArgs& G, const float* __restrict__ verts,
                                              const float* __restrict__ fnorm, uint32_t ti,
                                              float (&a)[kTriAttr]) {
    if (G.textured || G.hasVN) {
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = verts[9 * (size_t)ti + i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) a[9 + i] = fnorm[3 * (size_t)ti + i];
    const size_t lt = ti - G.triOff;
    if (G.textured) {
#pragma unroll
        for (int i = 0; i < 6; ++i) a[12 + i] = G.uv[6 * lt + i];
    }
    if (G.hasVN) {
#pragma unroll
        for (int i = 0; i < 9; ++i) a[18 + i] = G.vn[9 * lt + i];
    }
}

// Colour / normal update of one voxel with new hits from group G; get(k, a)
// gives the attributes of the voxel's k-th slot triangle (slots in reference
// order).
template <class Get>
__device__ void colour_update(const GroupArgs& G, f3 voxCenter, uint32_t count, int nslots,
                              uint32_t& col, f3& nrm, const float* __restrict__ inv255,
                              Get&& get) {
    f4 color = unpack4(col, inv255);
    f3 normal = nrm;
    if (G.textured) {
        uint32_t triCount = count < 4u ? count : 4u;
        for (int k = 0; k < nslots; ++k) {
            float a[kTriAttr];
            get(k, a);
            f3 v[3] = {{a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8]}};
            f3 fn = {a[9], a[10], a[11]};
            float ax = fabsf(fn.x), ay = fabsf(fn.y), az = fabsf(fn.z);
            int axis = (ax > ay && ax > az) ? 0 : (ay > az ? 1 : 2);
            f2 bc = bary_axis(axis, voxCenter, v);
            f3 addn;
            if (G.hasVN) {
                addn = interp_n(bc, f3{a[18], a[19], a[20]}, f3{a[21], a[22], a[23]},
                                f3{a[24], a[25], a[26]});
            } else {
                addn = fn;
            }
            normal = add3(normal, addn);
            f2 uv1 = {a[12], a[13]}, uv2 = {a[14], a[15]}, uv3 = {a[16], a[17]};
            f2 Vv = {(uv2.x - uv1.x) * bc.y, (uv2.y - uv1.y) * bc.y};
            f2 Uu = {(uv3.x - uv1.x) * bc.x, (uv3.y - uv1.y) * bc.x};
            f2 tc = {uv1.x + Uu.x + Vv.x, uv1.y + Uu.y + Vv.y};
            f4 tx = tex2d(G, tc.x, tc.y, inv255);
            if ((double)tx.w <= 0.005) {
                --triCount;
            } else {
                tx.x *= tx.w;
                tx.y *= tx.w;
                tx.z *= tx.w;
                color.x += tx.x;
                color.y += tx.y;
                color.z += tx.z;
                color.w += tx.w;
            }
        }
        if (triCount > 0u) {
            float fc = (float)triCount;
            color.x /= fc;
            color.y /= fc;
            color.z /= fc;
            color.w = gmin(color.w, 1.0f);
            normal.x /= fc;
            normal.y /= fc;
            normal.z /= fc;
            normal = normalize3(normal);
        }
    } else {
        uint32_t triCount = count;
        for (int k = 0; k < nslots; ++k) {
            float a[kTriAttr];
            get(k, a);
            f3 fn = {a[9], a[10], a[11]};
            if (G.hasVN) {
                f3 v[3] = {{a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8]}};
                int axis = (fn.x > fn.y && fn.x > fn.z) ? 0 : (fn.y > fn.z ? 1 : 2);
                f2 bc = bary_axis(axis, voxCenter, v);
                f3 ni = interp_n(bc, f3{a[18], a[19], a[20]}, f3{a[21], a[22], a[23]},
                                 f3{a[24], a[25], a[26]});
                color.x += ni.x;
                color.y += ni.y;
                color.z += ni.z;
                color.w += 1.0f;
            } else {
                color.x += fn.x;
                color.y += fn.y;
                color.z += fn.z;
                color.w += 1.0f;
            }
        }
        float fc = (float)triCount;
        color.x /= fc;
        color.y /= fc;
        color.z /= fc;
        color.w = gmin(color.w, 1.0f);
        normal = normalize3(f3{color.x, color.y, color.z});
        if (color.x < 0.0f) color.x *= -1.0f;
        if (color.y < 0.0f) color.y *= -1.0f;
        if (color.z < 0.0f) color.z *= -1.0f;
    }
    col = to_color(color);
    nrm = normal;
}

// Edge-test constants of TriangleOverlapsVoxelOnAxis (VK:93-144) for one
// axis pair with compile-time axes: per edge i, the per-voxel test is
// dp + uvw[A2] * vpy + uvw[A1] * vpx >= 0.
template <int A1, int A2, int OT>
__device__ __forceinline__ void stage_pair(const float (&v)[3][3], const float (&n)[3],
                                           const float* pp, const float* dd, float* st) {
    const float ps = n[OT] >= 0.0f ? 1.0f : -1.0f;
    const float e[3][3] = {{v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]},
                           {v[2][0] - v[1][0], v[2][1] - v[1][1], v[2][2] - v[1][2]},
                           {v[0][0] - v[2][0], v[0][1] - v[2][1], v[0][2] - v[2][2]}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float enx = (-e[i][A2]) * ps, eny = e[i][A1] * ps;
        const float vpx = enx * dd[A1], vpy = eny * dd[A2];
        const float dpp = ((-enx) * v[i][A1] + (-eny) * v[i][A2]) + (enx * pp[A1] + eny * pp[A2]) +
                          gmax(0.0f, dd[A1] * enx) + gmax(0.0f, dd[A2] * eny);
        st[i * 3 + 0] = dpp;
        st[i * 3 + 1] = vpx;
        st[i * 3 + 2] = vpy;
    }
}

// One unit of the voxelize pass: up to kVoxTiles consecutive tiles (ui *
// kVoxTiles ...) of a group's work list of nwork tiles.  Returns the SAT
// tests it ran.  Called by every thread of a workgroup; its LDS is reused by
// the workgroup's next unit.
__device__ __forceinline__ uint32_t voxelize_unit(
    const VoxParams& P, const GroupArgs& G, int init, const uint4* __restrict__ work,
    uint32_t nwork, uint32_t ui, const float* __restrict__ verts, const float* __restrict__ fnorm,
    const uint32_t* __restrict__ bbox, const uint32_t* __restrict__ binTri,
    uint32_t* __restrict__ occ, uint32_t* __restrict__ hit, uint4* __restrict__ rec,
    uint32_t* __restrict__ recCount, uint64_t segCap) {
    constexpr int kV = kVoxTiles * kTileVox, kW = kVoxTiles * kTileWords;
    __shared__ uint32_t s_cnt[kV];
    __shared__ int32_t s_lo0[kV], s_lo1[kV], s_lo2[kV], s_hi[kV];
    __shared__ float s_tri[kStage * kTriStride];
    __shared__ uint32_t s_box[kStage];
    __shared__ uint32_t s_tid[kStage];
    __shared__ uint32_t s_off[kStage + 1];
    __shared__ uint32_t s_w[kBlock / 64];

    __shared__ uint32_t s_magic[65];   // ceil(65536 / d) for d = 1..64
    __shared__ uint32_t s_hitw[kW], s_hitoff[kW], s_occprev[kW];
    __shared__ uint32_t s_recBase;
    __shared__ uint4 s_org[kVoxTiles];   // per tile of the unit: origin x, y, z, active tile
    __shared__ uint2 s_bin[kVoxTiles];   // per tile of the unit: first bin entry, entries
    const int tid = threadIdx.x;
    if (tid >= 1 && tid <= 64) s_magic[tid] = (65536u + (uint32_t)tid - 1u) / (uint32_t)tid;
    uint32_t localTests = 0;
    const uint32_t w0 = ui * kVoxTiles;
    const uint32_t nt = nwork - w0 < (uint32_t)kVoxTiles ? nwork - w0 : (uint32_t)kVoxTiles;
    // work entries: (active tile, tile index, first bin entry, entries)
    uint32_t nb = 0;
#pragma unroll
    for (int k = 0; k < kVoxTiles; ++k)
        if ((uint32_t)k < nt) nb += work[w0 + k].w;
    if ((uint32_t)tid < nt) {
        const uint4 wk = work[w0 + tid];
        uint32_t tx, ty, tz;
        tile_xyz(wk.y, P.TG, tx, ty, tz);
        s_org[tid] = make_uint4(tx * kTile, ty * kTile, tz * kTile, wk.x);
        s_bin[tid] = make_uint2(wk.z, wk.w);
    }
    for (int v = tid; v < kV; v += kBlock) {
        s_cnt[v] = 0u;
        s_lo0[v] = s_lo1[v] = s_lo2[v] = 0x7fffffff;
        s_hi[v] = -1;
    }
    if (tid < kW) {
        const uint32_t k = (uint32_t)tid / kTileWords;
        s_occprev[tid] = (init || k >= nt) ? 0u
                                           : occ[(uint64_t)work[w0 + k].x * kTileWords + tid % kTileWords];
    }
    __syncthreads();

    for (uint32_t c0 = 0; c0 < nb; c0 += kStage) {
        const uint32_t m = (nb - c0) < (uint32_t)kStage ? (nb - c0) : (uint32_t)kStage;
        // Thread j < m takes triangle j of the stage: its clipped voxel box,
        // and its staged constants -- per axis pair XY, ZX, YZ the edge-test
        // constants of TriangleOverlapsVoxelOnAxis (VK:93-144), with the
        // plane-test constants of TrianglePlaneOverlapsVoxel (VK:17-53) --
        // so the per-voxel work needs 2 mul + 2 add per edge.  Bounds,
        // vertices and normal are loaded together (one dependent round trip
        // after the bin entry).
        uint32_t w = 0;
        if ((uint32_t)tid < m) {
            // entry e of the unit's bins, tile 0's first: tile sel, place eb
            uint32_t sel = 0, eb = c0 + (uint32_t)tid;
            uint2 bn = s_bin[0];
            while (eb >= bn.y && sel + 1u < nt) { eb -= bn.y; bn = s_bin[++sel]; }
            const uint4 og = s_org[sel];
            const uint32_t ox = og.x, oy = og.y, oz = og.z;
            const uint32_t t = binTri[bn.x + eb];
            const uint32_t bx = bbox[3 * (size_t)t], by = bbox[3 * (size_t)t + 1],
                           bz = bbox[3 * (size_t)t + 2];
            const float* q = verts + 9 * (size_t)t;
            const float v[3][3] = {{q[0], q[1], q[2]}, {q[3], q[4], q[5]}, {q[6], q[7], q[8]}};
            const float* fq = fnorm + 3 * (size_t)t;
            const float n[3] = {fq[0], fq[1], fq[2]};
            uint32_t c0[3], c1[3];
            const bool inReach = clip_box(G, bx, by, bz, c0, c1);
            int32_t x0 = (int32_t)c0[0] - (int32_t)ox, x1 = (int32_t)c1[0] - (int32_t)ox;
            int32_t y0 = (int32_t)c0[1] - (int32_t)oy, y1 = (int32_t)c1[1] - (int32_t)oy;
            int32_t z0 = (int32_t)c0[2] - (int32_t)oz, z1 = (int32_t)c1[2] - (int32_t)oz;
            if (!inReach) x1 = -1;
            x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0; z0 = z0 < 0 ? 0 : z0;
            x1 = x1 > 7 ? 7 : x1; y1 = y1 > 7 ? 7 : y1; z1 = z1 > 7 ? 7 : z1;
            if (x0 <= x1 && y0 <= y1 && z0 <= z1) {
                w = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1));
                s_box[tid] = (uint32_t)x0 | ((uint32_t)y0 << 3) | ((uint32_t)z0 << 6) |
                             ((uint32_t)(x1 - x0) << 9) | ((uint32_t)(y1 - y0) << 12) |
                             ((uint32_t)(z1 - z0) << 15) | (sel << 18);
            } else {
                s_box[tid] = 0u;
            }
            s_tid[tid] = t;
            float* st = &s_tri[tid * kTriStride];
            st[0] = n[0]; st[1] = n[1]; st[2] = n[2];
            const float D = dot3(f3{-n[0], -n[1], -n[2]}, f3{v[0][0], v[0][1], v[0][2]});
            st[3] = D;
            stage_pair<0, 1, 2>(v, n, P.offP, P.dP, st + 4);
            stage_pair<2, 0, 1>(v, n, P.offP, P.dP, st + 13);
            stage_pair<1, 2, 0>(v, n, P.offP, P.dP, st + 22);
            // Plane-test bounds: the 8 corner distances are dist1 plus the
            // sums of the negative / positive parts of n_i * dP_i; thr bounds
            // the fp32 error of any corner distance (256 unit roundoffs of the
            // largest |n . p| + |D| over the grid).
            float neg = 0.0f, pos = 0.0f, M = fabsf(D);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float o = n[k] * P.dP[k];
                neg += fminf(o, 0.0f);
                pos += fmaxf(o, 0.0f);
                M += fabsf(n[k]) * (fabsf(P.offP[k]) + (float)(P.S[k] + 1u) * fabsf(P.dP[k]));
            }
            st[31] = neg;
            st[32] = pos;
            st[33] = M * (1.0f / 65536.0f);
        }
        uint32_t total;
        uint32_t off = block_excl_scan(w, s_w, &total);   // barrier
        if ((uint32_t)tid < m) s_off[tid] = off;
        if (tid == 0) s_off[m] = total;
        __syncthreads();
        // Each thread takes a contiguous range of (triangle, voxel) items: one
        // search for its first triangle, then an incremental walk over the
        // clipped bounds (x fastest), moving to the next triangle at its end.
        const uint32_t per = (total + kBlock - 1) / kBlock;
        uint32_t item = (uint32_t)tid * per;
        const uint32_t itemEnd = item + per < total ? item + per : total;
        uint32_t j = 0, bx = 0, ex = 1, ey = 1, cx = 0, cy = 0, cz = 0, jEnd = 0;
        uint32_t ox = 0, oy = 0, oz = 0;   // origin of the current triangle's tile
        // The current triangle's plane-test constants, held in registers
        // while the walk stays on it (face normal, D, plane-bound offsets and
        // error bound); the edge-test constants are read from LDS, only for
        // voxels that pass the plane test.
        float tn0 = 0.0f, tn1 = 0.0f, tn2 = 0.0f, tD = 0.0f, tNeg = 0.0f, tPos = 0.0f, tThr = 0.0f;
        const float* te = s_tri;
        auto load_tri = [&](uint32_t jj) {
            const float* st = &s_tri[jj * kTriStride];
            tn0 = st[0]; tn1 = st[1]; tn2 = st[2]; tD = st[3];
            te = st + 4;
            tNeg = st[31]; tPos = st[32]; tThr = st[33];
        };
        if (item < itemEnd) {
            uint32_t lo = 0, hi = m;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_off[mid] <= item) lo = mid; else hi = mid;
            }
            j = lo;
            const uint32_t r = item - s_off[j];
            bx = s_box[j];
            ex = ((bx >> 9) & 7u) + 1u;
            ey = ((bx >> 12) & 7u) + 1u;
            { const uint4 og = s_org[bx >> 18]; ox = og.x; oy = og.y; oz = og.z; }
            // r < 512: (r * ceil(65536 / d)) >> 16 == r / d for d <= 64
            const uint32_t exy = ex * ey;
            cz = (r * s_magic[exy]) >> 16;
            const uint32_t rxy = r - cz * exy;
            cy = (rxy * s_magic[ex]) >> 16;
            cx = rxy - cy * ex;
            jEnd = s_off[j + 1];
            load_tri(j);
        }
        for (; item < itemEnd; ++item) {
            if (item == jEnd) {
                // next triangle with a non-empty clipped box
                do { ++j; jEnd = s_off[j + 1]; } while (jEnd == item);
                bx = s_box[j];
                ex = ((bx >> 9) & 7u) + 1u;
                ey = ((bx >> 12) & 7u) + 1u;
                { const uint4 og = s_org[bx >> 18]; ox = og.x; oy = og.y; oz = og.z; }
                cx = cy = cz = 0;
                load_tri(j);
            }
            const uint32_t lx = (bx & 7u) + cx;
            const uint32_t ly = ((bx >> 3) & 7u) + cy;
            const uint32_t lz = ((bx >> 6) & 7u) + cz;
            if (++cx == ex) { cx = 0; if (++cy == ey) { cy = 0; ++cz; } }
            const float uvw[3] = {(float)(ox + lx), (float)(oy + ly), (float)(oz + lz)};
            ++localTests;
            // TrianglePlaneOverlapsVoxel (VK:17-53).
            const f3 n = {tn0, tn1, tn2};
            const float D = tD;
            const f3 p0 = {P.offP[0] + uvw[0] * P.dP[0], P.offP[1] + uvw[1] * P.dP[1],
                           P.offP[2] + uvw[2] * P.dP[2]};
            const float dist1 = dot3(n, p0) + D;
            // Certain outcomes first: every corner on one side (no hit) or
            // corners on both sides beyond the error bound (hit).  Only the
            // band where a corner may round to the other side runs the
            // reference's 8-corner products.
            const float lo = dist1 + tNeg, hi = dist1 + tPos, thr = tThr;
            bool planeHit = false;
            if (lo < -thr && hi > thr) {
                planeHit = true;
            } else if (!(lo > thr || hi < -thr)) {
                const float dx = P.dP[0], dy = P.dP[1], dz = P.dP[2];
                const f3 cd[7] = {{dx, dy, dz}, {dx, dy, 0.0f}, {dx, 0.0f, 0.0f}, {0.0f, 0.0f, dz},
                                  {0.0f, dy, dz}, {dx, 0.0f, dz}, {0.0f, dy, 0.0f}};
#pragma unroll
                for (int c = 0; c < 7; ++c) {
                    f3 p1 = add3(p0, cd[c]);
                    float dist2 = dot3(n, p1) + D;
                    planeHit = planeHit || (dist1 * dist2 <= 0.0f);
                }
            }
            if (!planeHit) continue;
            // TriangleOverlapsVoxel (VK:146-170): XY, ZX, YZ.
            bool edgeHit = true;
            const int A1[3] = {0, 2, 1}, A2[3] = {1, 0, 2};
#pragma unroll
            for (int pr = 0; pr < 3; ++pr) {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int c = (pr * 3 + i) * 3;
                    float ot = te[c] + uvw[A2[pr]] * te[c + 2] + uvw[A1[pr]] * te[c + 1];
                    edgeHit = edgeHit && (ot >= 0.0f);
                }
            }
            if (!edgeHit) continue;
            const uint32_t vi = ((bx >> 18) << 9) | ((lz * 8u + ly) * 8u + lx);
            const int32_t t = (int32_t)s_tid[j];
            atomicAdd(&s_cnt[vi], 1u);
            // Three lowest indices by an atomic-min cascade: each level keeps
            // min(old, x) and passes max(old, x) on, so level i ends holding
            // the i-th smallest index whatever the interleaving.  Highest by
            // atomic max.
            int32_t x = t;
            int32_t old = atomicMin(&s_lo0[vi], x);
            x = old > x ? old : x;
            if (x != 0x7fffffff) {
                old = atomicMin(&s_lo1[vi], x);
                x = old > x ? old : x;
                if (x != 0x7fffffff) atomicMin(&s_lo2[vi], x);
            }
            atomicMax(&s_hi[vi], t);
        }
        __syncthreads();
    }

    // Hit words by wave ballot, then one record per hit voxel (tile, voxel,
    // count, slots in reference order) for the colour pass.
    for (int pass = 0; pass < kV / kBlock; ++pass) {
        const uint32_t v = (uint32_t)(pass * kBlock + tid);
        const uint64_t mask = __ballot(s_cnt[v] > 0u);
        if ((tid & 63) == 0) {
            const uint32_t hw = v >> 5;  // even word index
            s_hitw[hw] = (uint32_t)mask;
            s_hitw[hw + 1] = (uint32_t)(mask >> 32);
        }
    }
    __syncthreads();
    if (tid < kW) {
        const uint32_t w = s_hitw[tid];
        const uint32_t k = (uint32_t)tid / kTileWords;
        s_hitoff[tid] = (uint32_t)__popc(w);
        if (k < nt) {
            const uint64_t o = (uint64_t)s_org[k].w * kTileWords + tid % kTileWords;
            hit[o] = w;
            occ[o] |= w;
        }
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < kW; ++w) { const uint32_t c = s_hitoff[w]; s_hitoff[w] = acc; acc += c; }
        const uint32_t seg = ui % kRecSegs;
        s_recBase = acc ? (uint32_t)(seg * segCap) + atomicAdd(&recCount[seg], acc) : 0u;
    }
    __syncthreads();
    for (int pass = 0; pass < kV / kBlock; ++pass) {
        const uint32_t v = (uint32_t)(pass * kBlock + tid);
        const uint32_t w = s_hitw[v >> 5];
        if ((w >> (v & 31u)) & 1u) {
            const uint32_t k = v >> 9, lv = v & (kTileVox - 1u);
            const uint32_t below = w & ((1u << (v & 31u)) - 1u);
            const uint32_t h = s_hitoff[v >> 5] + (uint32_t)__popc(below);
            const uint32_t cnt = s_cnt[v];
            const uint32_t wasHit = (s_occprev[v >> 5] >> (v & 31u)) & 1u;
            int32_t sl[4] = {-1, -1, -1, -1};
            const uint32_t k3 = cnt < 3u ? cnt : 3u;
            if (k3 > 0u) sl[0] = s_lo0[v];
            if (k3 > 1u) sl[1] = s_lo1[v];
            if (k3 > 2u) sl[2] = s_lo2[v];
            const uint32_t ns = k3 + (cnt >= 4u ? 1u : 0u);
            if (cnt >= 4u) sl[3] = s_hi[v];
            const size_t r = (size_t)(s_recBase + h) * 2;
            const uint4 og = s_org[k];
            const uint32_t X = og.x + (lv & 7u), Y = og.y + ((lv >> 3) & 7u), Z = og.z + (lv >> 6);
            rec[r] = make_uint4(og.w, lv | (ns << 9) | (wasHit << 12) | (Z << 16), cnt,
                                (uint32_t)sl[0]);
            rec[r + 1] = make_uint4((uint32_t)sl[1], (uint32_t)sl[2], (uint32_t)sl[3], X | (Y << 16));
        }
    }
    __syncthreads();   // LDS is reused by the next unit
    return localTests;
}

__global__ __launch_bounds__(kBlock, 4) void k_voxelize(
    VoxParams P, GroupArgs G, int g, int init, const uint4* __restrict__ work,
    const uint32_t* __restrict__ workCount, const int32_t* __restrict__ list,
    const float* __restrict__ verts, const float* __restrict__ fnorm,
    const uint32_t* __restrict__ bbox, const uint32_t* __restrict__ binStart,
    const uint32_t* __restrict__ binTri, uint32_t* __restrict__ occ, uint32_t* __restrict__ hit,
    uint4* __restrict__ rec, uint32_t* __restrict__ recCount, uint64_t segCap,
    unsigned long long* __restrict__ satCounter) {
    __shared__ uint32_t s_sum[kBlock / 64];
    const uint32_t nwork = *workCount;
    const uint32_t nunits = (nwork + kVoxTiles - 1) / kVoxTiles;
    uint32_t localTests = 0;
    // Persistent grid over units of kVoxTiles tiles that hold triangles of
    // this group: the fixed cost of a round (staging, scan, barriers, record
    // reservation) is paid once per unit.
    for (uint32_t ui = blockIdx.x; ui < nunits; ui += gridDim.x)
        localTests += voxelize_unit(P, G, init, work, nwork, ui, verts, fnorm, bbox, binTri, occ,
                                    hit, rec, recCount, segCap);
    if (satCounter) {
        uint32_t tot;
        block_excl_scan(localTests, s_sum, &tot);
        if (threadIdx.x == 0) atomicAdd(satCounter, (unsigned long long)tot);
    }
}

// The voxelize pass of many voxelizations in one persistent launch: the
// units of every job's group (job j owns units unitStart[j] ..
// unitStart[j+1] - 1) are handed out one at a time from a device queue, so
// workgroups that finish a small job's units go on with any other job's.
// Jobs are listed heaviest first, so the long units start early.
__global__ __launch_bounds__(kBlock, 4) void k_voxelize_batch(
    const VoxUnitJob* __restrict__ jobs, uint32_t njobs, const uint32_t* __restrict__ unitStart,
    uint32_t* __restrict__ queue) {
    __shared__ uint32_t s_u, s_j;
    const uint32_t total = unitStart[njobs];
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t u = atomicAdd(queue, 1u);
            uint32_t lo = 0, hi = njobs;   // last job whose first unit is <= u
            if (u < total)
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (unitStart[mid] <= u) lo = mid; else hi = mid;
                }
            s_u = u;
            s_j = lo;
        }
        __syncthreads();
        const uint32_t u = __builtin_amdgcn_readfirstlane(s_u);
        const uint32_t j = __builtin_amdgcn_readfirstlane(s_j);
        __syncthreads();   // s_u / s_j are written again only after every thread read them
        if (u >= total) break;
        const VoxUnitJob& J = jobs[j];
        voxelize_unit(J.P, J.G, J.init, J.work, *J.workCount, u - unitStart[j], J.verts, J.fnorm,
                      J.bbox, J.binTri, J.occ, J.hit, J.rec, J.recCount, J.segCap);
    }
}

// Units per job (from the work counts the jobs' group_tiles passes wrote),
// as a prefix, and the queue reset: one workgroup.
__global__ __launch_bounds__(kBlock) void k_voxelize_batch_units(const VoxUnitJob* __restrict__ jobs,
                                                                 uint32_t njobs,
                                                                 uint32_t* __restrict__ unitStart,
                                                                 uint32_t* __restrict__ queue) {
    __shared__ uint32_t s_w[kBlock / 64];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) { s_carry = 0; *queue = 0; }
    __syncthreads();
    for (uint32_t b = 0; b < njobs; b += kBlock) {
        const uint32_t j = b + threadIdx.x;
        const uint32_t n = j < njobs ? (*jobs[j].workCount + kVoxTiles - 1) / kVoxTiles : 0u;
        uint32_t tot;
        const uint32_t off = block_excl_scan(n, s_w, &tot);
        if (j < njobs) unitStart[j] = s_carry + off;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) unitStart[njobs] = s_carry;
}

// Colour pass over the hit records of one group: ComputeColorsAndNormals /
// ComputeColorsAsNormals (VoxelizerKernels.cu 689-1083) for each voxel with
// new hits, one thread per voxel (grid-stride).
// Records are taken kBlock at a time per workgroup; the distinct triangles
// their slots name are loaded once into LDS (a voxel's slot triangles are
// mostly shared with its neighbours' records), through a hash table of
// triangle indices; a slot whose triangle did not get a place there reads
// global memory. Either way the values are the same.
constexpr uint32_t kCKeys = 512;          // hash table entries
constexpr uint32_t kCTris = 224;          // staged triangles per chunk
constexpr uint32_t kTexLds = 1024;        // texels: a group image this small is read from LDS
constexpr uint32_t kCStride = kTriAttr + 1;
constexpr uint32_t kKeyEmpty = 0xffffffffu;
__global__ __launch_bounds__(kBlock) void k_colour(
    VoxParams P, GroupArgs G, int init, const uint4* __restrict__ rec,
    const uint32_t* __restrict__ recCount, uint64_t segCap, const int32_t* __restrict__ list,
    const float* __restrict__ verts, const float* __restrict__ fnorm, uint32_t* __restrict__ col,
    uint32_t* __restrict__ qn, float* __restrict__ nrm, const uint32_t* __restrict__ fillw) {
    __shared__ float s_inv255[256];
    __shared__ uint32_t s_pre[kRecSegs + 1];   // record segments: prefix of their counts
    __shared__ uint32_t s_key[kCKeys];         // triangle index, or kKeyEmpty
    __shared__ uint32_t s_idx[kCKeys];         // its staged place (>= kCTris: not staged)
    __shared__ uint32_t s_nt;
    __shared__ float s_dat[kCTris * kCStride];
    __shared__ uint32_t s_tex[kTexLds];
    const int tid = threadIdx.x;
    fill_inv255(s_inv255);
    // a small image (the usual height ramp) is read from LDS
    GroupArgs Gl = G;
    const uint64_t texels = (uint64_t)(G.w > 0 ? G.w : 0) * (uint64_t)(G.h > 0 ? G.h : 0);
    if (G.textured && G.img && texels <= kTexLds) {
        for (uint32_t k = tid; k < texels; k += kBlock)
            s_tex[k] = reinterpret_cast<const uint32_t*>(G.img)[k];
        Gl.img = reinterpret_cast<const uint8_t*>(s_tex);
    }
    for (uint32_t k = tid; k < kCKeys; k += kBlock) s_key[k] = kKeyEmpty;
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < kRecSegs; ++k) {
            s_pre[k] = acc;
            acc += recCount[k];
        }
        s_pre[kRecSegs] = acc;
        s_nt = 0u;
    }
    __syncthreads();
    const uint32_t n = s_pre[kRecSegs];
    // Staging pays for groups that read vertices (textured, or with vertex
    // normals); one that reads face normals only goes to global memory.
    const bool stage = G.textured || G.hasVN;   // uniform
    // the loop bound is the same for the whole workgroup (barriers inside)
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + (uint32_t)tid;
        const bool act = i < n;
        uint4 r0 = make_uint4(0u, 0u, 0u, 0u), r1 = make_uint4(0u, 0u, 0u, 0u);
        if (act) {
            uint32_t lo = 0, hi = kRecSegs;   // segment of the i-th record
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= i) lo = mid;
                else hi = mid;
            }
            const uint64_t ri = lo * segCap + (i - s_pre[lo]);
            r0 = rec[2 * ri];
            r1 = rec[2 * ri + 1];
        }
        const uint32_t ns = act ? (r0.y >> 9) & 7u : 0u;
        const uint32_t slots[4] = {r0.w, r1.x, r1.y, r1.z};
        // 1. a table entry per distinct slot triangle; the thread that
        // inserts one stages its attributes
        uint32_t hpos[4] = {kCKeys, kCKeys, kCKeys, kCKeys};
        uint32_t own = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (stage && (uint32_t)k < ns) {
                const uint32_t t = slots[k];
                uint32_t h = (t * 0x9E3779B1u) >> 23;   // 9 bits
                for (int probe = 0; probe < 8; ++probe) {
                    const uint32_t prev = atomicCAS(&s_key[h], kKeyEmpty, t);
                    if (prev == kKeyEmpty) { own |= 1u << k; hpos[k] = h; break; }
                    if (prev == t) { hpos[k] = h; break; }
                    h = (h + 1u) & (kCKeys - 1u);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if ((own >> k) & 1u) {
                const uint32_t d = atomicAdd(&s_nt, 1u);
                s_idx[hpos[k]] = d;
                if (d < kCTris) {
                    float a[kTriAttr];
                    load_tri_attr(G, verts, fnorm, slots[k], a);
                    float* sd = s_dat + d * kCStride;
#pragma unroll
                    for (int q = 0; q < 12; ++q) sd[q] = a[q];
                    if (G.textured) {
#pragma unroll
                        for (int q = 12; q < 18; ++q) sd[q] = a[q];
                    }
                    if (G.hasVN) {
#pragma unroll
                        for (int q = 18; q < kTriAttr; ++q) sd[q] = a[q];
                    }
                }
            }
        }
        if (stage) __syncthreads();
        // 2. the voxel's colour and normals
        if (act) {
            const uint64_t a = r0.x;
            const uint32_t v = r0.y & 511u, wasHit = (r0.y >> 12) & 1u;
            const uint32_t cnt = r0.z;
            const uint64_t o = a * kTileVox + v;
            // the voxel's value so far: stored when an earlier group hit it or a
            // terrain fill reached it, zero otherwise
            const bool prior = !init && (wasHit || ((fillw[a * kTileWords + (v >> 5)] >> (v & 31u)) & 1u));
            uint32_t c = prior ? col[o] : 0u;
            f3 nv = prior ? f3{nrm[(a * 3 + 0) * kTileVox + v], nrm[(a * 3 + 1) * kTileVox + v],
                               nrm[(a * 3 + 2) * kTileVox + v]}
                          : f3{0.0f, 0.0f, 0.0f};
            const uint32_t X = r1.w & 0xffffu, Y = r1.w >> 16, Z = r0.y >> 16;
            // the chunk's origin in float (voxOffsetOrigin, Voxelizer.cu
            // 2317-2319), then the voxel's place in the chunk (VK:715-719);
            // with one chunk the origin is offP itself
            const uint32_t cx = X - X % P.C[0], cy = Y - Y % P.C[1], cz = Z - Z % P.C[2];
            const f3 org = {P.offP[0] + (float)cx * P.dP[0], P.offP[1] + (float)cy * P.dP[1],
                            P.offP[2] + (float)cz * P.dP[2]};
            const f3 cpos = {org.x + (float)(X - cx) * P.dP[0], org.y + (float)(Y - cy) * P.dP[1],
                             org.z + (float)(Z - cz) * P.dP[2]};
            const f3 voxCenter = add3(cpos, muls3(f3{P.dP[0], P.dP[1], P.dP[2]}, 0.5f));
            colour_update(Gl, voxCenter, wasHit + cnt, (int)ns, c, nv, s_inv255,
                          [&](int k, float (&at)[kTriAttr]) {
                              const uint32_t hk = hpos[k];
                              const uint32_t d = hk < kCKeys ? s_idx[hk] : kCTris;
                              if (d < kCTris) {
                                  const float* sd = s_dat + d * kCStride;
#pragma unroll
                                  for (int q = 0; q < 12; ++q) at[q] = sd[q];
                                  if (G.textured) {
#pragma unroll
                                      for (int q = 12; q < 18; ++q) at[q] = sd[q];
                                  }
                                  if (G.hasVN) {
#pragma unroll
                                      for (int q = 18; q < kTriAttr; ++q) at[q] = sd[q];
                                  }
                              } else {
                                  load_tri_attr(G, verts, fnorm, slots[k], at);
                              }
                          });
            col[o] = c;
            qn[o] = quant_normal(nv);
            nrm[(a * 3 + 0) * kTileVox + v] = nv.x;
            nrm[(a * 3 + 1) * kTileVox + v] = nv.y;
            nrm[(a * 3 + 2) * kTileVox + v] = nv.z;
        }
        if (!stage) continue;
        __syncthreads();
        // 3. empty the table for the next chunk
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if ((own >> k) & 1u) s_key[hpos[k]] = kKeyEmpty;
        if (tid == 0) s_nt = 0u;
        __syncthreads();
    }
}

// Tiles holding triangles of group g, appended in any order (the tiles are
// independent, so the order does not change any result), each with what
// the voxelize pass needs first: (active tile, tile index, first bin entry,
// entries), so a workgroup starts a tile with one load.
__global__ void k_group_tiles(VoxParams P, int g, uint64_t nActive, const int32_t* __restrict__ list,
                              const uint32_t* __restrict__ binStart, uint4* __restrict__ work,
                              uint32_t* __restrict__ workCount) {
    const uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t tlin = 0, b0 = 0, b1 = 0;
    if (a < nActive) {
        tlin = (uint32_t)list[a];
        const uint64_t key = (uint64_t)g * P.numTiles + tlin;
        b0 = binStart[key];
        b1 = binStart[key + 1];
    }
    // one atomic per wave on the shared counter; lanes take consecutive slots
    const bool take = a < nActive && b1 > b0;
    const uint64_t mask = __ballot(take);
    if (!mask) return;
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(workCount, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (take)
        work[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] =
            make_uint4((uint32_t)a, tlin, b0, b1 - b0);
}

// ---------------------------------------------------------------------------
// fill: terrain fill (VoxelizerKernels.cu 886-917).  A voxel without new hits
// takes colour and normal of the nearest voxel above (at most 19) that has
// new hits in this group.  Per column of the tile a 32-bit mask of hits over
// this tile and the three above gives the nearest source by one bit scan.
// For the first group the voxelize / colour passes write hit voxels only, so
// here every other voxel gets its fill value or zero.
// One wave per active tile, kBlock / 64 tiles per workgroup; lane c owns the
// voxel column c = (x, y) of the tile and walks it top down, so a run of
// filled voxels under one hit voxel reads that source once. Lane q * 16 + w
// loads hit word w of the tile q tiles above (q = 0: the tile itself), one
// load for the 64 words of the four tiles; shuffles give each lane its
// column's hit bits over 32 layers. Voxels neither hit nor filled are not
// written (they read as zero through LevelMap::defOcc / defFill); the filled
// ones are recorded in fillw.
constexpr int kFillWaves = kBlock / 64;
__global__ __launch_bounds__(kBlock) void k_fill(VoxParams P, int g, int init, uint64_t nActive,
                                                 const int32_t* __restrict__ list,
                                                 const int32_t* __restrict__ tileIdx,
                                                 const uint32_t* __restrict__ binStart,
                                                 uint32_t* __restrict__ col,
                                                 uint32_t* __restrict__ qn,
                                                 float* __restrict__ nrm,
                                                 const uint32_t* __restrict__ hit,
                                                 uint32_t* __restrict__ fillw) {
    static_assert((kFillTiles + 1) * kTileWords == 64, "one hit word per lane");
    const int lane = (int)(threadIdx.x & 63u);
    const uint64_t a = (uint64_t)blockIdx.x * kFillWaves + (threadIdx.x >> 6);
    if (a >= nActive) return;   // whole waves
    const uint64_t tlin = (uint64_t)list[a];
    uint32_t tx_, ty_, tz;
    tile_xyz(tlin, P.TG, tx_, ty_, tz);
    const uint64_t slab = (uint64_t)P.TG[0] * P.TG[1];
    const uint32_t q = (uint32_t)lane >> 4, wd = (uint32_t)lane & 15u;
    int32_t idx = -1;
    if (q == 0) idx = (int32_t)a;
    else if (tz + q < P.TG[2]) idx = tileIdx[tlin + (uint64_t)q * slab];
    // the voxelize pass wrote hit words for the tiles holding triangles of
    // this group; every other tile has no hit of it
    uint32_t hw = 0u;
    if (idx >= 0) {
        const uint64_t key = (uint64_t)g * P.numTiles + tlin + (uint64_t)q * slab;
        const uint32_t w = hit[(uint64_t)idx * kTileWords + wd];   // beside the bin loads
        hw = binStart[key + 1] > binStart[key] ? w : 0u;
    }
    int32_t sidx[kFillTiles + 1];
#pragma unroll
    for (int k = 0; k <= kFillTiles; ++k) sidx[k] = __shfl(idx, k * kTileWords);
    // hit bits of column c, bit k * 8 + z for layer z of the tile k above
    const uint32_t c = (uint32_t)lane;
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k <= kFillTiles; ++k)
#pragma unroll
        for (int lz = 0; lz < kTile; ++lz) {
            const uint32_t word = (uint32_t)__shfl((int)hw, k * kTileWords + lz * 2 + (int)(c >> 5));
            m |= ((word >> (c & 31u)) & 1u) << (k * kTile + lz);
        }
    uint32_t lastZ = 0xffffffffu, sc = 0u, sq = 0u;
    float sn0 = 0.0f, sn1 = 0.0f, sn2 = 0.0f;
    // the fill stops at the chunk's z = 0 (the fill of a voxel only reads
    // voxels above it in the same chunk along z)
    const uint32_t zt = tz * kTile;
#pragma unroll
    for (int lz = kTile - 1; lz >= 0; --lz) {
        const uint32_t v = ((uint32_t)lz << 6) | c;
        const uint32_t zg = zt + (uint32_t)lz;
        const uint32_t zLim = (zg / P.C[2] + 1u) * P.C[2];   // first z of the next chunk
        const uint32_t reach = min(zLim - zg - 1u, (uint32_t)kFillDepth);
        const uint32_t above = (m >> (lz + 1)) & ((1u << reach) - 1u);
        const bool filled = !((m >> lz) & 1u) && above != 0u;
        if (filled) {
            const uint32_t zz = (uint32_t)lz + (uint32_t)__ffs(above);
            if (zz != lastZ) {   // a new source: read it once
                const uint32_t qq = zz >> 3, v2 = ((zz & 7u) << 6) | c;
                int32_t s_ = sidx[0];
#pragma unroll
                for (int k = 1; k <= kFillTiles; ++k) s_ = qq == (uint32_t)k ? sidx[k] : s_;
                const uint64_t sa = (uint64_t)s_;
                sc = col[sa * kTileVox + v2];
                sq = qn[sa * kTileVox + v2];
                sn0 = nrm[(sa * 3 + 0) * kTileVox + v2];
                sn1 = nrm[(sa * 3 + 1) * kTileVox + v2];
                sn2 = nrm[(sa * 3 + 2) * kTileVox + v2];
                lastZ = zz;
            }
            col[a * kTileVox + v] = sc;
            qn[a * kTileVox + v] = sq;
            nrm[(a * 3 + 0) * kTileVox + v] = sn0;
            nrm[(a * 3 + 1) * kTileVox + v] = sn1;
            nrm[(a * 3 + 2) * kTileVox + v] = sn2;
        }
        const uint64_t fb = __ballot(filled);
        if (lane == 0) {
            const uint64_t w = a * kTileWords + (uint64_t)lz * 2u;
            if (init) {
                fillw[w] = (uint32_t)fb;
                fillw[w + 1] = (uint32_t)(fb >> 32);
            } else {
                fillw[w] |= (uint32_t)fb;
                fillw[w + 1] |= (uint32_t)(fb >> 32);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// mip: ComputeVoxelMipMap8 (VoxelizerKernels.cu 1520-1663).
// With several chunks, voxel X of chunk i (X = i * Cd + x) reads source
// voxels i * Cs + 2x and i * Cs + 2x + 1 (Cd, Cs: the chunk's extent at the
// two levels), i.e. 2X + i * (Cs - 2 Cd).  Where Cd is a multiple of the
// tile edge, i * (Cs - 2 Cd) a multiple of two tiles and every source below
// the source level's end, a tile's sources are the 2x2x2 tiles at 2t
// shifted by that offset, with the same parities, and the filter is this
// one; shift.cd = 0 on an axis means one chunk (no shift).
__global__ __launch_bounds__(kBlock) void k_mip(LevelMap src, LevelMap dst, MipShift shift,
                                                const int32_t* __restrict__ listDst,
                                                const uint32_t* __restrict__ colSrc,
                                                const float* __restrict__ nrmSrc,
                                                uint32_t* __restrict__ colDst,
                                                uint32_t* __restrict__ qnDst,
                                                float* __restrict__ nrmDst) {
    const int tid = threadIdx.x;
    const uint64_t b = blockIdx.x;
    const uint64_t tlin = (uint64_t)listDst[b];
    uint32_t tx, ty, tz;
    tile_xyz(tlin, dst.TG, tx, ty, tz);
    // source offsets of this tile's chunk (multiples of 16 voxels)
    const uint32_t ox = shift.cd[0] ? (tx * kTile) / shift.cd[0] * shift.off[0] : 0u;
    const uint32_t oy = shift.cd[1] ? (ty * kTile) / shift.cd[1] * shift.off[1] : 0u;
    const uint32_t oz = shift.cd[2] ? (tz * kTile) / shift.cd[2] * shift.off[2] : 0u;
    __shared__ float s_inv255[256];
    // stored-voxel words of the (up to) 8 source tiles under this tile, when
    // the source is level 0; child c = (z & 1) * 4 + (y & 1) * 2 + (x & 1)
    __shared__ uint32_t s_def[8][kTileWords];
    fill_inv255(s_inv255);
    if (src.defOcc && tid < 8 * kTileWords) {
        const uint32_t c = (uint32_t)tid / kTileWords, w = (uint32_t)tid % kTileWords;
        const uint32_t sx = 2 * tx + (ox >> 3) + (c & 1u), sy = 2 * ty + (oy >> 3) + ((c >> 1) & 1u),
                       sz = 2 * tz + (oz >> 3) + (c >> 2);
        uint32_t d = 0u;
        if (sx < src.TG[0] && sy < src.TG[1] && sz < src.TG[2]) {
            const int32_t ca = src.tileIdx[((uint64_t)sz * src.TG[1] + sy) * src.TG[0] + sx];
            if (ca >= 0) {
                const uint64_t o = (uint64_t)ca * kTileWords + w;
                d = src.defOcc[o] | src.defFill[o];
            }
        }
        s_def[c][w] = d;
    }
    __syncthreads();
    // both passes unrolled: the source loads of the second are independent
    // of the first pass's stores (distinct arrays) and can go out early
#pragma unroll
    for (int pass = 0; pass < kTileVox / kBlock; ++pass) {
        const uint32_t v = (uint32_t)(pass * kBlock + tid);
        const uint32_t X = tx * kTile + (v & 7u), Y = ty * kTile + ((v >> 3) & 7u),
                       Z = tz * kTile + (v >> 6);
        uint32_t outC = 0u;
        f3 outN = {0.0f, 0.0f, 0.0f};
        if (X < dst.dims[0] && Y < dst.dims[1] && Z < dst.dims[2]) {
            const uint32_t bx = 2 * X + ox, by = 2 * Y + oy, bz = 2 * Z + oz;
            const uint32_t x1 = min(bx + 1, src.dims[0] - 1), y1 = min(by + 1, src.dims[1] - 1),
                           z1 = min(bz + 1, src.dims[2] - 1);
            const uint64_t st = ((uint64_t)(bz >> 3) * src.TG[1] + (by >> 3)) * src.TG[0] + (bx >> 3);
            const int32_t sa = src.tileIdx[st];
            if (sa >= 0) {
                // bx is even, so (bx, bx + 1) sit together at an 8-byte
                // aligned place of the tile: one 8-byte load per (z, y) row
                // and plane. When bx + 1 is past the level's edge the pair's
                // first voxel stands for both (x1 == bx).
                const uint32_t ys[2] = {by & 7u, y1 & 7u}, zs[2] = {bz & 7u, z1 & 7u};
                const bool xdup = x1 == bx;
                const uint64_t cb = (uint64_t)sa * kTileVox;
                const uint64_t nb0 = (uint64_t)sa * 3 * kTileVox;
                uint2 cp[4];
                float2 np[3][4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t row = (zs[j >> 1] * 8u + ys[j & 1]) * 8u + (bx & 7u);
                    // stored voxels of the pair (level 0 keeps only hit or
                    // filled voxels; the others read as zero)
                    uint32_t dm = 3u;
                    if (src.defOcc) {
                        const uint32_t ch = (((bz >> 3) & 1u) << 2) | (((by >> 3) & 1u) << 1) |
                                            ((bx >> 3) & 1u);
                        dm = (s_def[ch][row >> 5] >> (row & 31u)) & 3u;
                    }
                    cp[j] = make_uint2(0u, 0u);
                    np[0][j] = np[1][j] = np[2][j] = make_float2(0.0f, 0.0f);
                    if (dm) {
                        cp[j] = *reinterpret_cast<const uint2*>(colSrc + cb + row);
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            np[c][j] = *reinterpret_cast<const float2*>(nrmSrc + nb0 + c * kTileVox + row);
                        if (!(dm & 1u)) {
                            cp[j].x = 0u;
#pragma unroll
                            for (int c = 0; c < 3; ++c) np[c][j].x = 0.0f;
                        }
                        if (!(dm & 2u)) {
                            cp[j].y = 0u;
#pragma unroll
                            for (int c = 0; c < 3; ++c) np[c][j].y = 0.0f;
                        }
                    }
                    if (xdup) {
                        cp[j].y = cp[j].x;
#pragma unroll
                        for (int c = 0; c < 3; ++c) np[c][j].y = np[c][j].x;
                    }
                }
                // the sum in the order (z, y, x), x fastest
                f4 avg = unpack4(cp[0].x, s_inv255);
                f3 an = {np[0][0].x, np[1][0].x, np[2][0].x};
#pragma unroll
                for (int i = 1; i < 8; ++i) {
                    const int j = i >> 1;
                    f4 cur = unpack4((i & 1) ? cp[j].y : cp[j].x, s_inv255);
                    avg.x += cur.x; avg.y += cur.y; avg.z += cur.z; avg.w += cur.w;
                    an = add3(an, (i & 1) ? f3{np[0][j].y, np[1][j].y, np[2][j].y}
                                          : f3{np[0][j].x, np[1][j].x, np[2][j].x});
                }
                const float oneOverEight = 1.0f / 8.0f;
                avg.w *= oneOverEight;
                avg.x *= oneOverEight;
                avg.y *= oneOverEight;
                avg.z *= oneOverEight;
                an = muls3(an, oneOverEight);
                float len2 = dot3(an, an);
                if (len2 > 0.0f) {
                    float l = sqrtf(len2);
                    an = {an.x / l, an.y / l, an.z / l};
                } else {
                    an = {0.0f, 0.0f, 0.0f};
                }
                outC = to_color(avg);
                outN = an;
            }
        }
        colDst[b * kTileVox + v] = outC;
        qnDst[b * kTileVox + v] = quant_normal(outN);
        nrmDst[(b * 3 + 0) * kTileVox + v] = outN.x;
        nrmDst[(b * 3 + 1) * kTileVox + v] = outN.y;
        nrmDst[(b * 3 + 2) * kTileVox + v] = outN.z;
    }
}

// Measurement only (timing runs): the stored source voxels k_mip reads.
__global__ void k_mip_source_count(LevelMap src, LevelMap dst, uint64_t nActive,
                                   const int32_t* __restrict__ listDst,
                                   unsigned long long* __restrict__ out) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long n = 0;
    if (b < nActive) {
        uint32_t tx, ty, tz;
        tile_xyz((uint64_t)listDst[b], dst.TG, tx, ty, tz);
        for (uint32_t c = 0; c < 8u; ++c) {
            const uint32_t sx = 2 * tx + (c & 1u), sy = 2 * ty + ((c >> 1) & 1u), sz = 2 * tz + (c >> 2);
            if (sx >= src.TG[0] || sy >= src.TG[1] || sz >= src.TG[2]) continue;
            const int32_t ca = src.tileIdx[((uint64_t)sz * src.TG[1] + sy) * src.TG[0] + sx];
            if (ca < 0) continue;
            if (!src.defOcc) {
                n += kTileVox;
                continue;
            }
            for (int w = 0; w < kTileWords; ++w) {
                const uint64_t o = (uint64_t)ca * kTileWords + w;
                n += __popc(src.defOcc[o] | src.defFill[o]);
            }
        }
    }
    if (n) atomicAdd(out, n);
}

__device__ __forceinline__ int64_t tile_of(const LevelMap& lm, uint32_t X, uint32_t Y, uint32_t Z) {
    const uint64_t t = ((uint64_t)(Z >> 3) * lm.TG[1] + (Y >> 3)) * lm.TG[0] + (X >> 3);
    return lm.tileIdx[t];
}
__device__ __forceinline__ uint32_t vox_of(uint32_t X, uint32_t Y, uint32_t Z) {
    return ((Z & 7u) * 8u + (Y & 7u)) * 8u + (X & 7u);
}
// Whether voxel v of active tile a holds a stored value (always, except on
// level 0: hit by some group or reached by a terrain fill).
__device__ __forceinline__ bool stored(const LevelMap& lm, uint64_t a, uint32_t v) {
    if (!lm.defOcc) return true;
    const uint64_t w = a * kTileWords + (v >> 5);
    return (((lm.defOcc[w] | lm.defFill[w]) >> (v & 31u)) & 1u) != 0u;
}

// ---------------------------------------------------------------------------
// Node typing: ComputeOctTreeNodeConstColor + ComputeOctTreeNodeType with the
// border handlers (VoxelizerKernels.cu 1680-2380).  A node is NON-CONST when
// any voxel of its stored brick (interior + 1-voxel border, x < xLimit: the
// reference's two-launch x split, Voxelizer.cu 1081-1120, can leave the last
// x block unexamined) differs from the node's sample, the voxel at the
// brick's first interior corner.  Split in two passes:
//   node_sample  per node: sample colour, and NON-CONST already when the
//                sample is non-zero and the brick reaches an inactive tile
//                (whose voxels are all zero);
//   node_compare per active tile: every voxel is read once and compared with
//                the samples of the (at most 8) nodes whose bricks hold it.
__device__ __forceinline__ void node_sample_at(const LevelMap& lm, const uint32_t* __restrict__ col,
                                               uint32_t Rd, uint32_t Bx, uint32_t By, uint32_t Bz,
                                               uint32_t bx, uint32_t by, uint32_t bz,
                                               uint32_t xLimit, uint32_t* __restrict__ cc,
                                               uint32_t* __restrict__ flag, uint64_t n) {
    uint32_t nx, ny, nz;
    node_xyz(n, (uint32_t)(__ffs(Rd) - 1), nx, ny, nz);
    const uint32_t sx = nx * Bx + bx, sy = ny * By + by, sz = nz * Bz + bz;
    const int64_t sa = tile_of(lm, sx, sy, sz);
    const uint32_t sv = vox_of(sx, sy, sz);
    uint32_t sample = 0u;
    if (sa >= 0) {
        const uint32_t cv = col[(uint64_t)sa * kTileVox + sv];   // beside the mask load
        sample = stored(lm, (uint64_t)sa, sv) ? cv : 0u;
    }
    uint32_t f = 0u;
    if (sample != 0u) {
        const uint32_t r0x = sx - 1, r0y = sy - 1, r0z = sz - 1;
        uint32_t r1x = sx + Bx;
        if (r1x > xLimit - 1) r1x = xLimit - 1;
        const uint32_t r1y = sy + By, r1z = sz + Bz;
        if (r0x <= r1x) {
            // every tile of the brick's box, the loads independent of each
            // other (no early exit), several in flight
            const uint32_t tx0 = r0x >> 3, tx1 = r1x >> 3, ty0 = r0y >> 3, ty1 = r1y >> 3;
            const uint32_t nt = (tx1 - tx0 + 1) * (ty1 - ty0 + 1) * ((r1z >> 3) - (r0z >> 3) + 1);
            uint32_t tx = tx0, ty = ty0, tz = r0z >> 3;
#pragma unroll 4
            for (uint32_t k = 0; k < nt; ++k) {
                const uint64_t t = ((uint64_t)tz * lm.TG[1] + ty) * lm.TG[0] + tx;
                f |= lm.tileIdx[t] < 0 ? 1u : 0u;
                if (++tx > tx1) {
                    tx = tx0;
                    if (++ty > ty1) { ty = ty0; ++tz; }
                }
            }
        }
    }
    cc[n] = sample;
    flag[n] = f;
}
// The samples of every depth in one launch: thread n is node n - start[d] of
// depth d for start[d] <= n < start[d + 1].
__global__ void k_node_sample_all(SampleDepths S) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= S.start[S.L]) return;
    int d = 0;
    while (d + 1 < S.L && n >= S.start[d + 1]) ++d;
    node_sample_at(S.lm[d], S.col[d], S.Rd[d], S.B[0], S.B[1], S.B[2], S.border[d][0],
                   S.border[d][1], S.border[d][2], S.xLimit[d], S.cc[d], S.flag[d],
                   n - S.start[d]);
}

// nodes n whose stored brick [n*B + b - 1, n*B + b + B] holds coordinate X
__device__ __forceinline__ void node_span(uint32_t X, uint32_t B, uint32_t b, uint32_t Rd,
                                          int32_t& lo, int32_t& hi) {
    const int32_t x = (int32_t)X - (int32_t)b;   // voxel relative to interior start
    // n*B - 1 <= x <= n*B + B  <=>  (x - B)/B <= n <= (x + 1)/B
    int32_t l = x - (int32_t)B;
    lo = l >= 0 ? (l + (int32_t)B - 1) / (int32_t)B : -((-l) / (int32_t)B);
    const int32_t h1 = x + 1;
    hi = h1 >= 0 ? h1 / (int32_t)B : -((-h1 + (int32_t)B - 1) / (int32_t)B);
    if (lo < 0) lo = 0;
    if (hi > (int32_t)Rd - 1) hi = (int32_t)Rd - 1;
}

// One wave per active tile, kBlock / 64 tiles per workgroup: lane l holds the
// voxels (l & 7, l >> 3, z) for z = 0..7 in registers (eight coalesced
// 256-byte loads). The wave then walks the nodes whose stored bricks reach
// the tile (at most 8 for B >= 8): each node's sample is one wave-uniform
// load, compared with the lane's voxels inside that brick, and one ballot
// decides the node's flag. No LDS, no barrier.
constexpr int kNcTiles = kBlock / 64;
__global__ __launch_bounds__(kBlock) void k_node_compare(
    LevelMap lm, const int32_t* __restrict__ list, uint64_t nActive,
    const uint32_t* __restrict__ col, uint32_t Rd, uint32_t Bx, uint32_t By, uint32_t Bz,
    uint32_t bx, uint32_t by, uint32_t bz, uint32_t xLimit, const uint32_t* __restrict__ cc,
    uint32_t* __restrict__ flag) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint64_t a = (uint64_t)blockIdx.x * kNcTiles + (threadIdx.x >> 6);
    if (a >= nActive) return;   // whole waves
    uint32_t tx, ty, tz;
    tile_xyz((uint64_t)list[a], lm.TG, tx, ty, tz);
    const uint32_t B3[3] = {Bx, By, Bz}, b3[3] = {bx, by, bz};
    const uint32_t o3[3] = {tx * kTile, ty * kTile, tz * kTile};
    // nodes whose bricks reach the tile, per axis
    int32_t blo[3], bhi[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        int32_t lo, hi;
        node_span(o3[ax], B3[ax], b3[ax], Rd, blo[ax], hi);
        node_span(o3[ax] + kTile - 1, B3[ax], b3[ax], Rd, lo, bhi[ax]);
    }
    if (blo[0] > bhi[0] || blo[1] > bhi[1] || blo[2] > bhi[2]) return;
    const uint32_t X = o3[0] + ((uint32_t)lane & 7u), Y = o3[1] + ((uint32_t)lane >> 3);
    const bool colOk = X < xLimit && X < lm.dims[0] && Y < lm.dims[1];
    // z layers inside the level
    uint32_t zIn = 0u;
    uint32_t c[kTile];
    const uint32_t* tc = col + a * kTileVox + (uint64_t)lane;
    // stored-voxel words of the tile (level 0), lane w < 16 holding word w
    uint32_t dw = 0xffffffffu;
    if (lm.defOcc && lane < kTileWords)
        dw = lm.defOcc[a * kTileWords + lane] | lm.defFill[a * kTileWords + lane];
#pragma unroll
    for (int lz = 0; lz < kTile; ++lz) {
        bool st = true;
        if (lm.defOcc)
            st = (((uint32_t)__shfl((int)dw, lz * 2 + (lane >> 5)) >> (lane & 31)) & 1u) != 0u;
        const uint32_t cv = tc[lz * 64];   // every voxel of the tile is allocated
        c[lz] = st ? cv : 0u;
        if (o3[2] + (uint32_t)lz < lm.dims[2]) zIn |= 1u << lz;
    }
    for (int32_t nz = blo[2]; nz <= bhi[2]; ++nz) {
        // stored brick of node n on an axis: [n*B + b - 1, n*B + b + B]
        const int32_t z0 = nz * (int32_t)Bz + (int32_t)bz - 1 - (int32_t)o3[2];
        const int32_t z1 = z0 + (int32_t)Bz + 1;
        uint32_t zm = 0u;
#pragma unroll
        for (int lz = 0; lz < kTile; ++lz)
            if (lz >= z0 && lz <= z1) zm |= 1u << lz;
        zm &= zIn;
        if (!zm) continue;
        for (int32_t ny = blo[1]; ny <= bhi[1]; ++ny) {
            const int32_t y0 = ny * (int32_t)By + (int32_t)by - 1;
            const bool inY = (int32_t)Y >= y0 && (int32_t)Y <= y0 + (int32_t)By + 1;
            for (int32_t nx = blo[0]; nx <= bhi[0]; ++nx) {
                const int32_t x0 = nx * (int32_t)Bx + (int32_t)bx - 1;
                const bool inX = (int32_t)X >= x0 && (int32_t)X <= x0 + (int32_t)Bx + 1;
                const uint64_t n = ((uint64_t)nz * Rd + ny) * Rd + nx;
                const uint32_t s = cc[n];
                uint32_t ne = 0u;
#pragma unroll
                for (int lz = 0; lz < kTile; ++lz) ne |= (c[lz] != s ? 1u : 0u) << lz;
                const bool d = colOk && inX && inY && (ne & zm) != 0u;
                if (__ballot(d) && lane == 0) flag[n] = 1u;
            }
        }
    }
}

// out[d] = excl[offs[d + 1]] - excl[offs[d]]: the flags set in each range of
// an exclusive scan's input.
__global__ void k_range_counts(const uint32_t* __restrict__ excl, RangeOffs offs, int n,
                               uint32_t* __restrict__ out) {
    const int d = (int)threadIdx.x;
    if (d < n) out[d] = excl[offs.v[d + 1]] - excl[offs.v[d]];
}

__global__ void k_brick_assign_all(uint64_t allNodes, const uint32_t* __restrict__ flag,
                                   const uint32_t* __restrict__ excl, DepthOut D) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= allNodes) return;
    int d = 0;
    while (d + 1 < D.L && n >= D.off[d + 1]) ++d;
    const uint64_t ln = n - D.off[d];
    const uint32_t f = flag[n];
    D.types[d][ln] = f ? 1 : 0;
    if (f) {
        const uint32_t b = excl[n] - excl[D.off[d]];
        D.bidx[d][ln] = b;
        D.brickNode[d][b] = (uint32_t)ln;
    } else {
        D.bidx[d][ln] = 0xffffffffu;
    }
}


// brick gather: WriteBinaryBrick / WriteUByteBinaryBrick middle-brick case
// (VoxelBrickWriter.cu 725-818, 509-588): (B+2)^3 voxels, RGBA8 colour and
// RGB8 normal ((n+1)*0.5*255, truncated).

__device__ __forceinline__ void brick_gather_at(
    const LevelMap& lm, const uint32_t* __restrict__ col, const uint32_t* __restrict__ qn,
    const uint32_t* __restrict__ brickNode, uint32_t Rd, uint32_t Bx, uint32_t By, uint32_t Bz,
    uint32_t bx, uint32_t by, uint32_t bz, uint32_t* __restrict__ bcol,
    uint32_t* __restrict__ bnrm, uint64_t j) {
    // tiles under the brick, with their stored-voxel words when the level is
    // level 0: staged in LDS when they fit (at most 6 per axis, which any
    // B <= 39 meets); a larger brick reads the tile map and the words from
    // global memory for every voxel
    constexpr uint32_t kGatherTiles = 216;
    __shared__ int32_t s_tile[kGatherTiles];
    __shared__ uint32_t s_def[kGatherTiles * kTileWords];
    const uint64_t n = brickNode[j];
    uint32_t nx, ny, nz;
    node_xyz(n, (uint32_t)(__ffs(Rd) - 1), nx, ny, nz);
    const uint32_t r0x = nx * Bx + bx - 1, r0y = ny * By + by - 1, r0z = nz * Bz + bz - 1;
    const uint32_t ex = Bx + 2, ey = By + 2, ez = Bz + 2;
    const uint32_t t0x = r0x >> 3, t0y = r0y >> 3, t0z = r0z >> 3;
    const uint32_t ntx = ((r0x + ex - 1) >> 3) - t0x + 1, nty = ((r0y + ey - 1) >> 3) - t0y + 1,
                   ntz = ((r0z + ez - 1) >> 3) - t0z + 1;
    const uint32_t ntl = ntx * nty * ntz;
    const bool staged = ntl <= kGatherTiles;   // uniform over the workgroup
    for (uint32_t i = threadIdx.x; staged && i < ntl * kTileWords; i += kBlock) {
        const uint32_t tl = i / kTileWords, w = i % kTileWords;
        const uint32_t ix = t0x + tl % ntx, iy = t0y + (tl / ntx) % nty, iz = t0z + tl / (ntx * nty);
        const int32_t ta = lm.tileIdx[((uint64_t)iz * lm.TG[1] + iy) * lm.TG[0] + ix];
        if (w == 0) s_tile[tl] = ta;
        uint32_t d = 0xffffffffu;
        if (lm.defOcc) {
            d = 0u;
            if (ta >= 0) {
                const uint64_t o = (uint64_t)ta * kTileWords + w;
                d = lm.defOcc[o] | lm.defFill[o];
            }
        }
        s_def[i] = d;
    }
    __syncthreads();
    const uint32_t nvox = ex * ey * ez, exy = ex * ey;
    const float invx = 1.0f / (float)ex, invxy = 1.0f / (float)exy;
    // kGatherBatch voxels per thread at a time: their loads are all issued
    // before the first store, so they are in flight together.
    constexpr uint32_t kGatherBatch = 16;
    for (uint32_t i0 = threadIdx.x; i0 < nvox; i0 += kGatherBatch * kBlock) {
        uint32_t c[kGatherBatch], q[kGatherBatch];
#pragma unroll
        for (uint32_t u = 0; u < kGatherBatch; ++u) {
            const uint32_t i = i0 + u * kBlock;
            c[u] = 0u;
            q[u] = kZeroNormalQ;
            if (i < nvox) {
                const uint32_t z = udiv_small(i, exy, invxy);
                const uint32_t rxy = i - z * exy;
                const uint32_t y = udiv_small(rxy, ex, invx);
                const uint32_t x = rxy - y * ex;
                const uint32_t X = r0x + x, Y = r0y + y, Z = r0z + z;
                const uint32_t v = vox_of(X, Y, Z);
                int32_t ta;
                bool def;
                if (staged) {
                    const uint32_t tl =
                        (((Z >> 3) - t0z) * nty + ((Y >> 3) - t0y)) * ntx + ((X >> 3) - t0x);
                    ta = s_tile[tl];
                    def = ((s_def[tl * kTileWords + (v >> 5)] >> (v & 31u)) & 1u) != 0u;
                } else {
                    ta = (int32_t)tile_of(lm, X, Y, Z);
                    def = ta >= 0 && stored(lm, (uint64_t)ta, v);
                }
                if (ta >= 0 && def) {
                    const uint64_t o = (uint64_t)ta * kTileVox + v;
                    c[u] = col[o];
                    q[u] = qn[o];
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kGatherBatch; ++u) {
            const uint32_t i = i0 + u * kBlock;
            if (i < nvox) {
                bcol[j * nvox + i] = c[u];
                bnrm[j * nvox + i] = q[u];
            }
        }
    }
}

// The bricks of every depth in one launch: workgroup w gathers brick
// w - start[e] of entry e for start[e] <= w < start[e + 1].
__global__ __launch_bounds__(kBlock) void k_brick_gather_all(GatherDepths G) {
    const uint64_t w = blockIdx.x;
    int e = 0;
    while (e + 1 < G.n && w >= G.start[e + 1]) ++e;
    brick_gather_at(G.lm[e], G.col[e], G.qn[e], G.brickNode[e], G.Rd[e], G.B[0], G.B[1], G.B[2],
                    G.border[e][0], G.border[e][1], G.border[e][2], G.bcol[e], G.bnrm[e],
                    w - G.start[e]);
}

// CollapseConstantBranches (Voxelizer.cu 701-774), one level at a time,
// leaves first.
__global__ void k_collapse(uint32_t Rd, int childIsLeaf, uint8_t* __restrict__ types,
                           const uint8_t* __restrict__ childTypes) {
    uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nn = (uint64_t)Rd * Rd * Rd;
    if (n >= nn) return;
    if (types[n] != 0) return;
    uint32_t x, y, z;
    node_xyz(n, (uint32_t)(__ffs(Rd) - 1), x, y, z);
    const uint32_t Rc = Rd * 2;
    bool all = true;
    for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t cx = 2 * x + (c & 1u), cy = 2 * y + ((c >> 1) & 1u), cz = 2 * z + (c >> 2);
        const uint8_t t = childTypes[((uint64_t)cz * Rc + cy) * Rc + cx];
        all = all && (childIsLeaf ? (t == 0) : (t == 2));
    }
    if (all) types[n] = 2;
}

__global__ __launch_bounds__(kBlock) void k_occupancy(VoxParams P, const int32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ occ,
                                                      uint32_t* __restrict__ bitmap) {
    const uint64_t a = blockIdx.x;
    const uint64_t tlin = (uint64_t)list[a];
    uint32_t tx, ty, tz;
    tile_xyz(tlin, P.TG, tx, ty, tz);
    if ((P.S[0] & 7u) == 0u) {
        // a tile row of 8 voxels is one byte of the bitmap
        const uint32_t r = threadIdx.x;
        if (r < 64u) {
            const uint32_t byte = (occ[a * kTileWords + (r >> 2)] >> ((r & 3u) * 8u)) & 255u;
            if (byte) {
                const uint64_t Y = ty * kTile + (r & 7u), Z = tz * kTile + (r >> 3);
                const uint64_t i = (Z * P.S[1] + Y) * P.S[0] + (uint64_t)tx * kTile;
                reinterpret_cast<uint8_t*>(bitmap)[i >> 3] = (uint8_t)byte;
            }
        }
        return;
    }
    for (int pass = 0; pass < kTileVox / kBlock; ++pass) {
        const uint32_t v = (uint32_t)(pass * kBlock + threadIdx.x);
        if (!((occ[a * kTileWords + (v >> 5)] >> (v & 31u)) & 1u)) continue;
        const uint64_t X = tx * kTile + (v & 7u), Y = ty * kTile + ((v >> 3) & 7u),
                       Z = tz * kTile + (v >> 6);
        const uint64_t i = (Z * P.S[1] + Y) * P.S[0] + X;
        atomicOr(&bitmap[i >> 5], 1u << (i & 31u));
    }
}

// The whole occupancy bitmap when a voxel row of S[0] bits is a whole number
// of 32-bit words: one workgroup per (y, z) row of tiles, whose tile indices
// it keeps in LDS, writes every word of the row's 64 voxel rows once (zeros
// included), so the bitmap needs no clearing first.
__global__ __launch_bounds__(kBlock) void k_occupancy_rows(VoxParams P,
                                                           const int32_t* __restrict__ tileIdx,
                                                           const uint32_t* __restrict__ occ,
                                                           uint32_t* __restrict__ bitmap) {
    extern __shared__ int32_t s_rowIdx[];   // P.TG[0] entries
    const uint32_t ty = blockIdx.x % P.TG[1], tz = blockIdx.x / P.TG[1];
    const uint64_t rowBase = ((uint64_t)tz * P.TG[1] + ty) * P.TG[0];
    for (uint32_t i = threadIdx.x; i < P.TG[0]; i += kBlock) s_rowIdx[i] = tileIdx[rowBase + i];
    __syncthreads();
    const uint32_t wpr = P.S[0] >> 5;   // words per voxel row; 4 tiles per word
    const uint32_t ny = P.S[1] - ty * kTile < (uint32_t)kTile ? P.S[1] - ty * kTile : (uint32_t)kTile;
    const uint32_t nz = P.S[2] - tz * kTile < (uint32_t)kTile ? P.S[2] - tz * kTile : (uint32_t)kTile;
    for (uint32_t k = threadIdx.x; k < 64u * wpr; k += kBlock) {
        const uint32_t r = k / wpr, w = k - r * wpr;   // tile row r = (z & 7) * 8 + (y & 7)
        const uint32_t ly = r & 7u, lz = r >> 3;
        if (ly >= ny || lz >= nz) continue;
        uint32_t word = 0u;
#pragma unroll
        for (uint32_t b = 0; b < 4u; ++b) {
            const int32_t a = s_rowIdx[w * 4u + b];
            if (a >= 0)
                word |= ((occ[(uint64_t)a * kTileWords + (r >> 2)] >> ((r & 3u) * 8u)) & 255u)
                        << (8u * b);
        }
        const uint64_t Y = (uint64_t)ty * kTile + ly, Z = (uint64_t)tz * kTile + lz;
        bitmap[(Z * P.S[1] + Y) * wpr + w] = word;
    }
}

// The same when a voxel row is a whole number of 16-byte runs (S[0] a
// multiple of 128): a thread takes 16 tiles along x and one occupancy word of
// each (four voxel rows of the tiles: word g of a tile holds rows 4g .. 4g+3,
// 8 bits each), loads the 16 words together and writes the four rows' 16
// bytes each with one store.
constexpr uint32_t kOccQuadBlock = 128;
__global__ __launch_bounds__(kOccQuadBlock) void k_occupancy_quads(
    VoxParams P, const int32_t* __restrict__ tileIdx, const uint32_t* __restrict__ occ,
    uint32_t* __restrict__ bitmap) {
    extern __shared__ int32_t s_rowIdx[];   // P.TG[0] entries
    const uint32_t ty = blockIdx.x % P.TG[1], tz = blockIdx.x / P.TG[1];
    const uint64_t rowBase = ((uint64_t)tz * P.TG[1] + ty) * P.TG[0];
    for (uint32_t i = threadIdx.x; i < P.TG[0]; i += kOccQuadBlock)
        s_rowIdx[i] = tileIdx[rowBase + i];
    __syncthreads();
    const uint32_t wpr = P.S[0] >> 5, qpr = wpr >> 2;   // words, 16-byte runs per voxel row
    const uint32_t ny = P.S[1] - ty * kTile < (uint32_t)kTile ? P.S[1] - ty * kTile : (uint32_t)kTile;
    const uint32_t nz = P.S[2] - tz * kTile < (uint32_t)kTile ? P.S[2] - tz * kTile : (uint32_t)kTile;
    for (uint32_t k = threadIdx.x; k < kTileWords * qpr; k += kOccQuadBlock) {
        const uint32_t g = k / qpr, qx = k - g * qpr;   // occupancy word g, tiles 16qx ..
        const uint32_t lz = g >> 1, ly0 = (g & 1u) * 4u;
        if (lz >= nz || ly0 >= ny) continue;
        uint32_t raw[16];
#pragma unroll
        for (uint32_t tt = 0; tt < 16u; ++tt) {
            const int32_t a = s_rowIdx[qx * 16u + tt];
            raw[tt] = a >= 0 ? occ[(uint64_t)a * kTileWords + g] : 0u;
        }
        const uint64_t Z = (uint64_t)tz * kTile + lz;
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            if (ly0 + j >= ny) break;
            uint32_t wd[4];
#pragma unroll
            for (uint32_t b = 0; b < 4u; ++b) {
                uint32_t x = 0u;
#pragma unroll
                for (uint32_t tt = 0; tt < 4u; ++tt)
                    x |= ((raw[4u * b + tt] >> (8u * j)) & 255u) << (8u * tt);
                wd[b] = x;
            }
            const uint64_t Y = (uint64_t)ty * kTile + ly0 + j;
            *reinterpret_cast<uint4*>(bitmap + (Z * P.S[1] + Y) * wpr + 4u * qx) =
                make_uint4(wd[0], wd[1], wd[2], wd[3]);
        }
    }
}

__global__ void k_dense_level(LevelMap lm, const uint32_t* __restrict__ col,
                              const float* __restrict__ nrm, uint32_t* __restrict__ outCol,
                              float* __restrict__ outNrm) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)lm.dims[0] * lm.dims[1] * lm.dims[2];
    if (i >= n) return;
    const uint32_t X = (uint32_t)(i % lm.dims[0]), Y = (uint32_t)((i / lm.dims[0]) % lm.dims[1]),
                   Z = (uint32_t)(i / ((uint64_t)lm.dims[0] * lm.dims[1]));
    const int64_t ta = tile_of(lm, X, Y, Z);
    uint32_t c = 0u;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    const uint32_t v = vox_of(X, Y, Z);
    if (ta >= 0 && stored(lm, (uint64_t)ta, v)) {
        c = col[(uint64_t)ta * kTileVox + v];
        nx = nrm[((uint64_t)ta * 3 + 0) * kTileVox + v];
        ny = nrm[((uint64_t)ta * 3 + 1) * kTileVox + v];
        nz = nrm[((uint64_t)ta * 3 + 2) * kTileVox + v];
    }
    outCol[i] = c;
    outNrm[3 * i] = nx;
    outNrm[3 * i + 1] = ny;
    outNrm[3 * i + 2] = nz;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan in one pass (decoupled look-back): each
// workgroup takes the next tile in launch order from a ticket counter, sums
// it, publishes the sum, then adds up its predecessors' published values
// (sums, or inclusive prefixes, which end the walk) and publishes its own
// inclusive prefix.  A predecessor holds an earlier ticket, so it is running
// or done and publishes its sum without waiting on anything: every wait ends.
//
// Scratch (zeroed once when allocated; each scan leaves it ready for the
// next): word 0-1 a 64-bit counter (call generation << 32 | next ticket),
// then one 64-bit status per tile (flag << 32 | value).  Flags carry the
// call's generation, so statuses of earlier calls read as not ready and
// nothing is cleared between calls; the holder of the last ticket starts the
// next generation.
constexpr int kScanItems = 32;
constexpr int kScanTile = kBlock * kScanItems;
// statuses each lane reads per look-back step (a window of 64 * 4 tiles)
constexpr int kScanLook = 4;
// bound on status polls (each with a short sleep): a hang guard only, since
// every awaited status is published without waiting
constexpr uint32_t kScanSpin = 1u << 22;

__device__ __forceinline__ uint64_t scan_status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void scan_status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void k_scan(const uint32_t* __restrict__ in, uint64_t n,
                                                 uint32_t nb, uint32_t* __restrict__ out,
                                                 uint32_t* __restrict__ clearIn,
                                                 unsigned long long* __restrict__ tmp,
                                                 uint32_t* __restrict__ err) {
    __shared__ uint32_t s_w[kBlock / 64];
    __shared__ unsigned long long s_tk;
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid == 0) s_tk = atomicAdd(tmp, 1ull);
    __syncthreads();
    const unsigned long long tk = s_tk;
    const uint32_t t = (uint32_t)tk;
    const uint32_t gen = (uint32_t)(tk >> 32) & 0x3fffffffu;   // flags stay below 2^31 + 1
    const uint32_t fAgg = 2u * gen + 1u, fPre = 2u * gen + 2u;
    uint64_t* status = reinterpret_cast<uint64_t*>(tmp + 1);

    // Wave w owns values [w * kScanTile / 4, (w + 1) * kScanTile / 4) of the
    // tile in kScanItems / 4 rows of 256; lane l holds values 4l .. 4l + 3 of
    // each row, so every load and store instruction of a wave covers 1 KB.
    const uint32_t wave = (uint32_t)tid >> 6;
    const uint64_t tileBase = (uint64_t)t * kScanTile;
    const uint64_t waveBase = tileBase + (uint64_t)wave * (kScanTile / 4);
    constexpr int kRows = kScanItems / 4;
    const bool full = tileBase + kScanTile <= n;
    uint4 vals[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint64_t i = waveBase + (uint64_t)k * 256u + (uint64_t)lane * 4u;
        if (full) {
            vals[k] = *reinterpret_cast<const uint4*>(in + i);
        } else {
            vals[k].x = i < n ? in[i] : 0u;
            vals[k].y = i + 1 < n ? in[i + 1] : 0u;
            vals[k].z = i + 2 < n ? in[i + 2] : 0u;
            vals[k].w = i + 3 < n ? in[i + 3] : 0u;
        }
    }
    // exclusive prefix of each lane's run of four within the wave's values
    uint32_t rowExcl[kRows];
    uint32_t waveSum = 0;
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint32_t sum = vals[k].x + vals[k].y + vals[k].z + vals[k].w;
        const uint32_t inc = wave_incl_scan(sum);
        rowExcl[k] = waveSum + inc - sum;
        waveSum += __shfl(inc, 63, 64);
    }
    if (lane == 0) s_w[wave] = waveSum;
    __syncthreads();
    uint32_t wavePre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t x = s_w[w];
        wavePre += (uint32_t)w < wave ? x : 0u;
        tot += x;
    }
    if (tid < 64) {
        uint32_t excl = 0;
        if (t == 0) {
            if (lane == 0) {
                scan_status_store(status, ((uint64_t)fPre << 32) | tot);
                // the last ticket starts the next generation at ticket 0
                if (nb == 1) scan_status_store(reinterpret_cast<uint64_t*>(tmp), (tk + (1ull << 32)) & ~0xffffffffull);
            }
        } else {
            if (lane == 0) scan_status_store(status + t, ((uint64_t)fAgg << 32) | tot);
            // the wave reads the 64 * kScanLook statuses before tile j0 + 1
            // at once, lane l those at distances kScanLook * l + q; positions
            // before tile 0 read as an inclusive prefix of 0
            int64_t j0 = (int64_t)t - 1;
            uint32_t polls = 0;
            while (true) {
                uint64_t st[kScanLook];
                while (true) {
                    bool ready = true;
#pragma unroll
                    for (int q = 0; q < kScanLook; ++q) {
                        const int64_t j = j0 - (int64_t)(kScanLook * lane + q);
                        st[q] = j >= 0 ? scan_status_load(status + j) : ((uint64_t)fPre << 32);
                        const uint32_t f = (uint32_t)(st[q] >> 32);
                        ready = ready && (f == fAgg || f == fPre);
                    }
                    if (__ballot(!ready) == 0ull || ++polls >= kScanSpin) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                // the nearest inclusive prefix ends the walk: the lanes
                // before its lane count in full, its lane up to it
                int firstPre = kScanLook;
#pragma unroll
                for (int q = kScanLook - 1; q >= 0; --q)
                    if ((uint32_t)(st[q] >> 32) == fPre) firstPre = q;
                const uint64_t pre = __ballot(firstPre < kScanLook);
                const int preLane = pre ? __ffsll((unsigned long long)pre) - 1 : 64;
                const int upto = lane < preLane ? kScanLook : (lane == preLane ? firstPre + 1 : 0);
                uint32_t v = 0u;
#pragma unroll
                for (int q = 0; q < kScanLook; ++q)
                    if (q < upto) v += (uint32_t)st[q];
                excl += __shfl(wave_incl_scan(v), 63, 64);
                if (pre) break;
                if (polls >= kScanSpin) {
                    // a predecessor never published: the prefix is wrong;
                    // report it instead of passing it on silently
                    if (lane == 0 && err) atomicOr(err, 1u);
                    break;
                }
                j0 -= 64 * kScanLook;
            }
            if (lane == 0) {
                scan_status_store(status + t, ((uint64_t)fPre << 32) | (excl + tot));
                if (t == nb - 1)
                    scan_status_store(reinterpret_cast<uint64_t*>(tmp), (tk + (1ull << 32)) & ~0xffffffffull);
            }
        }
        if (lane == 0) s_prefix = excl;
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
    // clearIn: the same words as in (read above)
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint64_t i = waveBase + (uint64_t)k * 256u + (uint64_t)lane * 4u;
        uint4 o;
        o.x = prefix + wavePre + rowExcl[k];
        o.y = o.x + vals[k].x;
        o.z = o.y + vals[k].y;
        o.w = o.z + vals[k].z;
        if (full) {
            *reinterpret_cast<uint4*>(out + i) = o;
            if (clearIn) *reinterpret_cast<uint4*>(clearIn + i) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            const uint32_t ov[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (i + e < n) {
                    out[i + e] = ov[e];
                    if (clearIn) clearIn[i + e] = 0u;
                }
            }
        }
    }
    if (t == nb - 1 && tid == 0) out[n] = prefix + tot;
}

// Order-independent 64-bit checksum of n words: the sum of a mixed
// (word, index) per word.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Several buffers hashed by one launch: segment s owns the blocks from
// segs[s].block0 up to the next segment's; each block adds its partial sum
// to the segment's output word.
__global__ void k_hash_segs(const HashSeg* __restrict__ segs, uint32_t nseg,
                            unsigned long long* __restrict__ outs) {
    __shared__ unsigned long long part[kBlock / 64];
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs[mid].block0 <= b) lo = mid;
        else hi = mid;
    }
    const HashSeg sg = segs[lo];
    const uint32_t nb = (lo + 1 < nseg ? segs[lo + 1].block0 : gridDim.x) - sg.block0;
    unsigned long long acc = 0ull;
    for (uint64_t i = (uint64_t)(b - sg.block0) * blockDim.x + threadIdx.x; i < sg.n;
         i += (uint64_t)nb * blockDim.x)
        acc += mix64(((unsigned long long)sg.p[i] << 32 | (unsigned long long)(i & 0xffffffffu)) ^
                     (sg.salt * 0x9E3779B97F4A7C15ull + (i >> 32)));
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0ull;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) t += part[w];
        atomicAdd(&outs[sg.out], t);
    }
}

// Rows idx[k] - base of src (w floats each) into row k of dst.
__global__ void k_gather_rows(const float* __restrict__ src, const uint32_t* __restrict__ idx,
                              uint64_t n, uint32_t base, uint32_t w, float* __restrict__ dst) {
    const uint64_t total = n * w;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = e / w;
        const uint32_t c = (uint32_t)(e - k * w);
        dst[e] = src[(uint64_t)(idx[k] - base) * w + c];
    }
}

// Order-preserving int image of a float, for atomic min / max.
__device__ __forceinline__ int32_t ord_f32(float f) {
    const int32_t i = __float_as_int(f);
    return i >= 0 ? i : (int32_t)(0x80000000u - (uint32_t)i) - 1;
}

// Bounds of n triangles' vertices (9 floats each): out[0..2] min, [3..5] max
// as ordered ints (out starts at INT_MAX / INT_MIN).
__global__ void k_tri_bounds(const float* __restrict__ v, uint64_t n, int32_t* __restrict__ out) {
    int32_t mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int32_t mx[3] = {(int32_t)0x80000000, (int32_t)0x80000000, (int32_t)0x80000000};
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        for (int k = 0; k < 9; ++k) {
            const int32_t o = ord_f32(v[9 * t + k]);
            const int a = k % 3;
            mn[a] = o < mn[a] ? o : mn[a];
            mx[a] = o > mx[a] ? o : mx[a];
        }
    for (int a = 0; a < 3; ++a)
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t m = __shfl_down(mn[a], o), M = __shfl_down(mx[a], o);
            mn[a] = m < mn[a] ? m : mn[a];
            mx[a] = M > mx[a] ? M : mx[a];
        }
    if ((threadIdx.x & 63u) == 0u)
        for (int a = 0; a < 3; ++a) {
            atomicMin(&out[a], mn[a]);
            atomicMax(&out[3 + a], mx[a]);
        }
}

// Plain copy of n 16-byte words (the reference point of traffic counters).
__global__ void k_copy_u128(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    // four coalesced 16-byte loads in flight per lane, then their stores
    const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i + 3 * T < n; i += 4 * T) {
        const uint4 a = src[i], b = src[i + T], c = src[i + 2 * T], d = src[i + 3 * T];
        dst[i] = a;
        dst[i + T] = b;
        dst[i + 2 * T] = c;
        dst[i + 3 * T] = d;
    }
    for (; i < n; i += T) dst[i] = src[i];
}

// The same copy over 256-byte blocks in a scattered order: 16 lanes per
// block, block i of the grid-stride walk copying block (i * odd) mod n.
__global__ void k_copy_scattered(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                 uint64_t nblocks) {
    const uint64_t lanes = nblocks * 16;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lanes;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = ((i >> 4) * 0x9E3779B1ull) & (nblocks - 1);
        const uint64_t w = b * 16 + (i & 15u);
        dst[w] = src[w];
    }
}

struct FillSegs {
    FillSeg s[4];
    int k;
};
// Grid-stride over the words of all segments, one after the other.
// Four words per thread: one 16-byte store where the words are 16-byte
// aligned and all in range, single-word stores otherwise.
__device__ __forceinline__ void fill4(uint32_t* p, uint64_t i, uint64_t n, uint32_t v) {
    if (i + 4 <= n && ((reinterpret_cast<uintptr_t>(p + i) & 15u) == 0u)) {
        *reinterpret_cast<uint4*>(p + i) = make_uint4(v, v, v, v);
    } else {
        for (uint64_t j = i; j < i + 4 && j < n; ++j) p[j] = v;
    }
}
// Segment k is filled in (n_k + 3) / 4 runs of four words.
__global__ void k_fill_segs(FillSegs F) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;; q += stride) {
        uint64_t r = q;
        int k = 0;
        while (k < F.k && r >= (F.s[k].n + 3) / 4) { r -= (F.s[k].n + 3) / 4; ++k; }
        if (k >= F.k) return;
        fill4(F.s[k].p, r * 4, F.s[k].n, F.s[k].v);
    }
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q * 4 < n; q += stride)
        fill4(p, q * 4, n, v);
}

// ---------------------------------------------------------------------------
// Several chunks (generateVoxelsAndOctTree with voxChunkDim smaller than the
// voxel space, Voxelizer.cu 2180-2814).  Level k is laid out chunk by chunk:
// chunk i of an axis holds voxels [i * Ck, (i + 1) * Ck) of its own mip chain
// (computeMipMaps per chunk, Voxelizer.cu 491-594), so level k voxel X of
// chunk i = X / Ck reads source voxels i * Cs + 2 (X - i * Ck) and the next
// one (clamped into the chunk).  Voxels past the level's extent read as zero.
__device__ __forceinline__ uint32_t level_col(const LevelMap& lm, const uint32_t* __restrict__ col,
                                              uint32_t X, uint32_t Y, uint32_t Z) {
    if (X >= lm.dims[0] || Y >= lm.dims[1] || Z >= lm.dims[2]) return 0u;
    const int64_t a = tile_of(lm, X, Y, Z);
    if (a < 0) return 0u;
    const uint32_t v = vox_of(X, Y, Z);
    return stored(lm, (uint64_t)a, v) ? col[(uint64_t)a * kTileVox + v] : 0u;
}
// level_col and level_nrm of one voxel with one tile lookup.
__device__ __forceinline__ uint32_t level_cn(const LevelMap& lm, const uint32_t* __restrict__ col,
                                             const float* __restrict__ nrm, uint32_t X, uint32_t Y,
                                             uint32_t Z, f3& n) {
    n = {0.0f, 0.0f, 0.0f};
    if (X >= lm.dims[0] || Y >= lm.dims[1] || Z >= lm.dims[2]) return 0u;
    const int64_t a = tile_of(lm, X, Y, Z);
    if (a < 0) return 0u;
    const uint32_t v = vox_of(X, Y, Z);
    if (!stored(lm, (uint64_t)a, v)) return 0u;
    const uint64_t b = (uint64_t)a * 3 * kTileVox + v;
    n = {nrm[b], nrm[b + kTileVox], nrm[b + 2 * kTileVox]};
    return col[(uint64_t)a * kTileVox + v];
}
// Source voxels of level voxel X along one axis.
__device__ __forceinline__ void chunk_src(uint32_t X, uint32_t Cd, uint32_t Cs, uint32_t& s0,
                                          uint32_t& s1) {
    const uint32_t i = X / Cd, x = X - i * Cd;
    s0 = i * Cs + 2u * x;
    s1 = i * Cs + min(2u * x + 1u, Cs - 1u);
}

__global__ void k_chunk_mark(LevelMap src, LevelMap dst, ChunkGeom cg, int ks, int kd,
                             uint32_t* __restrict__ markDst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dst.numTiles) return;
    uint32_t p[3];
    tile_xyz(t, dst.TG, p[0], p[1], p[2]);
    // per axis the distinct source tiles the tile's voxels read (at most 8:
    // a voxel's two sources share a tile)
    uint32_t st[3][kTile], ns[3] = {0u, 0u, 0u};
    for (int a = 0; a < 3; ++a) {
        for (uint32_t q = 0; q < (uint32_t)kTile; ++q) {
            const uint32_t X = p[a] * kTile + q;
            if (X >= dst.dims[a]) break;
            uint32_t s0, s1;
            chunk_src(X, cg.C[a] >> kd, cg.C[a] >> ks, s0, s1);
            const uint32_t tt = s0 >> 3;
            if (tt >= src.TG[a]) continue;
            if (ns[a] == 0u || st[a][ns[a] - 1u] != tt) st[a][ns[a]++] = tt;
        }
    }
    uint32_t m = 0u;
    for (uint32_t k = 0; k < ns[2] && !m; ++k)
        for (uint32_t j = 0; j < ns[1] && !m; ++j)
            for (uint32_t i = 0; i < ns[0]; ++i)
                if (src.tileIdx[((uint64_t)st[2][k] * src.TG[1] + st[1][j]) * src.TG[0] + st[0][i]] >= 0) {
                    m = 1u;
                    break;
                }
    markDst[t] = m;
}

// ComputeVoxelMipMap8 (VoxelizerKernels.cu 1520-1663) inside each chunk:
// the same sum order and arithmetic as k_mip, sources addressed per chunk.
__global__ __launch_bounds__(kBlock) void k_chunk_mip(LevelMap src, LevelMap dst, ChunkGeom cg,
                                                      int ks, int kd,
                                                      const int32_t* __restrict__ listDst,
                                                      const uint32_t* __restrict__ colSrc,
                                                      const float* __restrict__ nrmSrc,
                                                      uint32_t* __restrict__ colDst,
                                                      uint32_t* __restrict__ qnDst,
                                                      float* __restrict__ nrmDst) {
    const int tid = threadIdx.x;
    const uint64_t b = blockIdx.x;
    const uint64_t tlin = (uint64_t)listDst[b];
    uint32_t tx, ty, tz;
    tile_xyz(tlin, dst.TG, tx, ty, tz);
    __shared__ float s_inv255[256];
    fill_inv255(s_inv255);
    __syncthreads();
    for (int pass = 0; pass < kTileVox / kBlock; ++pass) {
        const uint32_t v = (uint32_t)(pass * kBlock + tid);
        const uint32_t X = tx * kTile + (v & 7u), Y = ty * kTile + ((v >> 3) & 7u),
                       Z = tz * kTile + (v >> 6);
        uint32_t outC = 0u;
        f3 outN = {0.0f, 0.0f, 0.0f};
        if (X < dst.dims[0] && Y < dst.dims[1] && Z < dst.dims[2]) {
            uint32_t xs[2], ys[2], zs[2];
            chunk_src(X, cg.C[0] >> kd, cg.C[0] >> ks, xs[0], xs[1]);
            chunk_src(Y, cg.C[1] >> kd, cg.C[1] >> ks, ys[0], ys[1]);
            chunk_src(Z, cg.C[2] >> kd, cg.C[2] >> ks, zs[0], zs[1]);
            // the sum in the order (z, y, x), x fastest
            f4 avg = {0.0f, 0.0f, 0.0f, 0.0f};
            f3 an = {0.0f, 0.0f, 0.0f};
            // the eight source voxels' values loaded before any is summed
            uint32_t cs[8];
            f3 ns[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                cs[i] = level_cn(src, colSrc, nrmSrc, xs[i & 1], ys[(i >> 1) & 1], zs[i >> 2],
                                 ns[i]);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const f4 cur = unpack4(cs[i], s_inv255);
                const f3 cn = ns[i];
                if (i == 0) {
                    avg = cur;
                    an = cn;
                } else {
                    avg.x += cur.x; avg.y += cur.y; avg.z += cur.z; avg.w += cur.w;
                    an = add3(an, cn);
                }
            }
            const float oneOverEight = 1.0f / 8.0f;
            avg.w *= oneOverEight;
            avg.x *= oneOverEight;
            avg.y *= oneOverEight;
            avg.z *= oneOverEight;
            an = muls3(an, oneOverEight);
            float len2 = dot3(an, an);
            if (len2 > 0.0f) {
                float l = sqrtf(len2);
                an = {an.x / l, an.y / l, an.z / l};
            } else {
                an = {0.0f, 0.0f, 0.0f};
            }
            outC = to_color(avg);
            outN = an;
        }
        colDst[b * kTileVox + v] = outC;
        qnDst[b * kTileVox + v] = quant_normal(outN);
        nrmDst[(b * 3 + 0) * kTileVox + v] = outN.x;
        nrmDst[(b * 3 + 1) * kTileVox + v] = outN.y;
        nrmDst[(b * 3 + 2) * kTileVox + v] = outN.z;
    }
}

// Voxels of a chunk the reference's typing launch reaches along one axis
// (Voxelizer.cu 1037-1120): tpb = the chunk's extent, cut back past the
// node grid's end (unsigned arithmetic as there), 8-thread blocks, and on x
// the split into two launches whose halved block count is rounded up only
// when odd.
__device__ __forceinline__ uint32_t chunk_typed_limit(uint32_t Ck, uint32_t o, uint32_t bo,
                                                      uint32_t full, bool xAxis) {
    uint32_t tpb = Ck;
    if (tpb + o - bo > full) tpb = full - (o - bo) + 2u;
    if (tpb == 0u) return 0u;
    const uint64_t th = tpb < 8u ? tpb : 8u;
    uint64_t nb = ((uint64_t)tpb + th - 1u) / th;
    uint64_t cover = nb * th;
    if (xAxis && tpb > 128u) {
        nb >>= 1;
        if (nb & 1u) nb += 1u;
        cover = 2ull * nb * th;
    }
    return cover < Ck ? (uint32_t)cover : Ck;
}

// ComputeOctTreeNodeConstColor + ComputeOctTreeNodeType per chunk
// (VoxelizerKernels.cu 1680-2380, Voxelizer.cu 931-1172), one wave per
// node: for every chunk the node's stored brick reaches, the chunk's sample
// (the brick's first interior voxel clamped into the chunk) against the
// brick's voxels in that chunk that the chunk's typing launch reaches; the
// node is NON-CONST when any differs (types OR-ed over chunks), and its
// colour is the sample of the last chunk in visit order.
//
// A brick window with no active tile under it reads zero everywhere (every
// sample lies in the window, so the colour is zero too): CONST, colour 0,
// without further reads.  Otherwise the window's voxel offsets u along
// each axis map to a chunk slot j (the chunk holding lo + u, counted from
// the window's first) and to whether that chunk's typing launch reaches
// them (kChunkTypedBit); the chunks' samples are read once per node (slot
// triple -> colour) and the wave stops at the first voxel that differs.
// Windows wider than kChunkAxis or with more than kChunkSamples slot
// triples read every voxel's sample directly.  The nodes of every depth in
// one launch, the coarse depths (few nodes, each a long walk) first, so
// their waves run beside the many waves of the finer depths.
constexpr uint32_t kChunkAxis = 66, kChunkSamples = 64, kChunkTypedBit = 1u << 31;
constexpr int kChunkWaves = kBlock / 64;
__global__ __launch_bounds__(kBlock) void k_chunk_node_type(SampleDepths S, ChunkGeom cg) {
    __shared__ uint32_t s_ax[kChunkWaves][3][kChunkAxis];
    __shared__ uint32_t s_sco[kChunkWaves][3][kChunkSamples];
    __shared__ uint32_t s_samp[kChunkWaves][kChunkSamples];
    const uint32_t lane = threadIdx.x & 63u;
    // the wave's index, and below its node and depth, in scalar registers
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t Bv[3] = {S.B[0], S.B[1], S.B[2]};
    const uint32_t e[3] = {Bv[0] + 2u, Bv[1] + 2u, Bv[2] + 2u};
    const uint64_t total = S.start[S.L];
    for (uint64_t g = (uint64_t)blockIdx.x * kChunkWaves + wv; g < total;
         g += (uint64_t)gridDim.x * kChunkWaves) {
        int dd = 0;   // the node's depth (uniform over the wave)
        while (g >= S.start[dd + 1]) ++dd;
        dd = __builtin_amdgcn_readfirstlane(dd);
        const uint64_t n = g - S.start[dd];
        const int k = S.L - 1 - dd;
        const uint32_t Rd = S.Rd[dd];
        const LevelMap& lm = S.lm[dd];
        const uint32_t* __restrict__ col = S.col[dd];
        uint32_t* __restrict__ cc = S.cc[dd];
        uint32_t* __restrict__ flag = S.flag[dd];
        const uint32_t Ck[3] = {cg.C[0] >> k, cg.C[1] >> k, cg.C[2] >> k},
                       bo[3] = {S.border[dd][0], S.border[dd][1], S.border[dd][2]};
        const uint32_t full[3] = {Rd * Bv[0], Rd * Bv[1], Rd * Bv[2]};
        const uint32_t ld = (uint32_t)(__ffs(Rd) - 1);
        uint32_t nd[3];
        node_xyz(n, ld, nd[0], nd[1], nd[2]);
        uint32_t lo[3], hi[3], c0[3], c1[3], sp[3], cov[3], m[3];
        bool inAny = true;
        for (int a = 0; a < 3; ++a) {
            sp[a] = nd[a] * Bv[a] + bo[a];
            lo[a] = sp[a] - 1u;
            hi[a] = sp[a] + Bv[a];
            cov[a] = cg.nc[a] * Ck[a];
            c0[a] = lo[a] / Ck[a];
            c1[a] = min(hi[a] / Ck[a], cg.nc[a] - 1u);
            m[a] = (lo[a] + e[a] - 1u) / Ck[a] - c0[a] + 1u;
            if (c0[a] >= cg.nc[a]) inAny = false;
        }
        // any active tile under the window
        bool any = false;
        if (inAny) {
            const uint32_t t0[3] = {lo[0] >> 3, lo[1] >> 3, lo[2] >> 3};
            uint32_t t1[3];
            for (int a = 0; a < 3; ++a) t1[a] = min(hi[a], cov[a] - 1u) >> 3;
            const uint32_t ntx = t1[0] - t0[0] + 1u, nty = t1[1] - t0[1] + 1u;
            const uint32_t ntl = ntx * nty * (t1[2] - t0[2] + 1u);
            bool mine = false;
            for (uint32_t q = lane; q < ntl; q += 64u) {
                const uint32_t qx = q % ntx, qy = (q / ntx) % nty, qz = q / (ntx * nty);
                const uint32_t tx = t0[0] + qx, ty = t0[1] + qy, tz = t0[2] + qz;
                if (tx < lm.TG[0] && ty < lm.TG[1] && tz < lm.TG[2] &&
                    lm.tileIdx[((uint64_t)tz * lm.TG[1] + ty) * lm.TG[0] + tx] >= 0)
                    mine = true;
            }
            any = __ballot(mine) != 0ull;
        }
        if (!any) {
            if (lane == 0) {
                cc[n] = 0u;
                flag[n] = 0u;
            }
            continue;
        }
        const bool tables = e[0] <= kChunkAxis && e[1] <= kChunkAxis && e[2] <= kChunkAxis &&
                            m[0] * m[1] * m[2] <= kChunkSamples;   // uniform
        bool diff = false;
        if (tables) {
            uint32_t (&ax)[3][kChunkAxis] = s_ax[wv];
            uint32_t (&sco)[3][kChunkSamples] = s_sco[wv];
            uint32_t* samp = s_samp[wv];
            for (uint32_t t = lane; t < e[0] + e[1] + e[2]; t += 64u) {
                const int a = t < e[0] ? 0 : (t < e[0] + e[1] ? 1 : 2);
                const uint32_t u = t - (a == 0 ? 0u : (a == 1 ? e[0] : e[0] + e[1]));
                const uint32_t P = lo[a] + u;
                const uint32_t ci = P / Ck[a], o = ci * Ck[a];
                const bool typed = P < cov[a] &&
                                   P - o < chunk_typed_limit(Ck[a], o, bo[a], full[a], a == 0);
                ax[a][u] = (typed ? kChunkTypedBit : 0u) | (ci - c0[a]);
            }
            for (uint32_t t = lane; t < m[0] + m[1] + m[2]; t += 64u) {
                const int a = t < m[0] ? 0 : (t < m[0] + m[1] ? 1 : 2);
                const uint32_t j = t - (a == 0 ? 0u : (a == 1 ? m[0] : m[0] + m[1]));
                const uint32_t o = (c0[a] + j) * Ck[a];
                sco[a][j] = min(max(sp[a], o), o + Ck[a] - 1u);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            // slot triples: the sample and, for the chunks the node's colour
            // is taken from (c0..c1), the visit rank
            const uint32_t mt = m[0] * m[1] * m[2];
            uint32_t myRank = 0u, myCol = 0u;
            bool ranked = false;
            for (uint32_t t = lane; t < mt; t += 64u) {
                const uint32_t j0 = t % m[0], j1 = (t / m[0]) % m[1], j2 = t / (m[0] * m[1]);
                const uint32_t c = level_col(lm, col, sco[0][j0], sco[1][j1], sco[2][j2]);
                samp[t] = c;
                if (c0[0] + j0 <= c1[0] && c0[1] + j1 <= c1[1] && c0[2] + j2 <= c1[2]) {
                    const uint32_t r = cg.rank[((uint64_t)(c0[2] + j2) * cg.nc[1] + (c0[1] + j1)) *
                                                   cg.nc[0] + (c0[0] + j0)];
                    if (!ranked || r > myRank) {
                        ranked = true;
                        myRank = r;
                        myCol = c;
                    }
                }
            }
            // the last chunk in visit order (ranks are distinct)
            uint32_t best = ranked ? myRank : 0u, bestCol = myCol;
            bool haveAny = ranked;
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t r2 = (uint32_t)__shfl_xor((int)best, off);
                const uint32_t c2 = (uint32_t)__shfl_xor((int)bestCol, off);
                const bool h2 = __shfl_xor((int)haveAny, off) != 0;
                if (h2 && (!haveAny || r2 > best)) {
                    best = r2;
                    bestCol = c2;
                    haveAny = true;
                }
            }
            if (lane == 0) cc[n] = haveAny ? bestCol : 0u;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            // voxel q = (w * e1 + v) * e0 + u; q / e0 and q / (e0 e1) by
            // multiply-high with ceil(2^32 / d), exact for q * d < 2^32
            // (windows of at most 66^3 voxels)
            const uint32_t e01 = e[0] * e[1], nv = e01 * e[2];
            const uint32_t inv0 = (uint32_t)((0x100000000ull + e[0] - 1u) / e[0]);
            const uint32_t inv01 = (uint32_t)((0x100000000ull + e01 - 1u) / e01);
            for (uint32_t b = 0; b < nv; b += 64u) {
                const uint32_t q = b + lane;
                bool d = false;
                if (q < nv) {
                    const uint32_t w = __umulhi(q, inv01), r = q - w * e01;
                    const uint32_t v = __umulhi(r, inv0), u = r - v * e[0];
                    const uint32_t ax0 = ax[0][u], ay = ax[1][v], az = ax[2][w];
                    if (ax0 & ay & az & kChunkTypedBit) {
                        const uint32_t j0 = ax0 & 0xffffu, j1 = ay & 0xffffu, j2 = az & 0xffffu;
                        const uint32_t scol = samp[(j2 * m[1] + j1) * m[0] + j0];
                        d = level_col(lm, col, lo[0] + u, lo[1] + v, lo[2] + w) != scol;
                    }
                }
                if (__ballot(d)) {
                    diff = true;
                    break;
                }
            }
            // the tables are rewritten by this wave's next node only after
            // every lane has left this one
            __builtin_amdgcn_wave_barrier();
        } else {
            // the constant colour: the last chunk in visit order
            if (lane == 0) {
                uint32_t best = 0u, bestRank = 0u;
                bool have = false;
                for (uint32_t kk = c0[2]; kk <= c1[2]; ++kk)
                    for (uint32_t j = c0[1]; j <= c1[1]; ++j)
                        for (uint32_t i = c0[0]; i <= c1[0]; ++i) {
                            const uint32_t r = cg.rank[((uint64_t)kk * cg.nc[1] + j) * cg.nc[0] + i];
                            if (!have || r > bestRank) {
                                have = true;
                                bestRank = r;
                                const uint32_t ci[3] = {i, j, kk};
                                uint32_t q[3];
                                for (int a = 0; a < 3; ++a)
                                    q[a] = min(max(sp[a], ci[a] * Ck[a]), ci[a] * Ck[a] + Ck[a] - 1u);
                                best = level_col(lm, col, q[0], q[1], q[2]);
                            }
                        }
                cc[n] = best;
            }
            const uint32_t nv = e[0] * e[1] * e[2];
            for (uint32_t b = 0; b < nv; b += 64u) {
                const uint32_t q = b + lane;
                bool d = false;
                if (q < nv) {
                    const uint32_t P3[3] = {lo[0] + q % e[0], lo[1] + (q / e[0]) % e[1],
                                            lo[2] + q / (e[0] * e[1])};
                    bool typed = true;
                    uint32_t s3[3];
                    for (int a = 0; a < 3; ++a) {
                        if (P3[a] >= cov[a]) { typed = false; break; }
                        const uint32_t ci = P3[a] / Ck[a], o = ci * Ck[a];
                        if (P3[a] - o >= chunk_typed_limit(Ck[a], o, bo[a], full[a], a == 0)) {
                            typed = false;
                            break;
                        }
                        s3[a] = min(max(sp[a], o), o + Ck[a] - 1u);
                    }
                    if (typed)
                        d = level_col(lm, col, P3[0], P3[1], P3[2]) !=
                            level_col(lm, col, s3[0], s3[1], s3[2]);
                }
                if (__ballot(d)) {
                    diff = true;
                    break;
                }
            }
        }
        if (lane == 0) flag[n] = diff ? 1u : 0u;
    }
}

// Brick keys: a NON-CONST node whose stored brick lies in processed chunks
// is written by the chunk that completes it (the brick writer's write
// order, Voxelizer.cu 1173-1262 and 2778, VoxelBrickWriter.cu 1131-1210):
// inside one chunk it is written while that chunk is typed; across chunks
// it is assembled and written after the last of them, with the other
// bricks completed there, in (z, y, x) order.
__global__ void k_chunk_brick_keys(const uint32_t* __restrict__ flag, ChunkGeom cg, int k,
                                   uint32_t Rd, uint32_t B0, uint32_t B1, uint32_t B2,
                                   uint32_t bo0, uint32_t bo1, uint32_t bo2,
                                   unsigned long long* __restrict__ keys,
                                   uint32_t* __restrict__ nodes, uint32_t* __restrict__ count) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nn = (uint64_t)Rd * Rd * Rd;
    if (n >= nn || flag[n] != 1u) return;
    const uint32_t Ck[3] = {cg.C[0] >> k, cg.C[1] >> k, cg.C[2] >> k}, Bv[3] = {B0, B1, B2},
                   bo[3] = {bo0, bo1, bo2};
    const uint32_t ld = (uint32_t)(__ffs(Rd) - 1);
    uint32_t nd[3];
    node_xyz(n, ld, nd[0], nd[1], nd[2]);
    uint32_t c0[3], c1[3];
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
        const uint32_t lo = nd[a] * Bv[a] + bo[a] - 1u, hi = nd[a] * Bv[a] + bo[a] + Bv[a];
        if (hi >= cg.nc[a] * Ck[a]) return;   // part of it never processed: no brick
        c0[a] = lo / Ck[a];
        c1[a] = hi / Ck[a];
        inside = inside && c0[a] == c1[a];
    }
    uint32_t last = 0u;
    for (uint32_t k = c0[2]; k <= c1[2]; ++k)
        for (uint32_t j = c0[1]; j <= c1[1]; ++j)
            for (uint32_t i = c0[0]; i <= c1[0]; ++i)
                last = max(last, cg.rank[((uint64_t)k * cg.nc[1] + j) * cg.nc[0] + i]);
    const uint32_t sh = 3u * ld;
    const uint32_t slot = atomicAdd(count, 1u);
    keys[slot] = ((unsigned long long)last << (sh + 1u)) | ((unsigned long long)(inside ? 0u : 1u) << sh) |
                 (unsigned long long)n;
    nodes[slot] = (uint32_t)n;
}

__global__ void k_chunk_types(uint64_t nn, const uint32_t* __restrict__ flag,
                              uint8_t* __restrict__ types, uint32_t* __restrict__ bidx) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    types[n] = (uint8_t)flag[n];
    bidx[n] = 0xffffffffu;
}
__global__ void k_chunk_bidx(uint32_t nb, const uint32_t* __restrict__ sorted,
                             uint32_t* __restrict__ bidx, uint32_t* __restrict__ brickNode) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const uint32_t n = sorted[j];
    bidx[n] = j;
    brickNode[j] = n;
}

inline unsigned grid_for(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// Fill kernel instead of hipMemsetAsync so the pipeline stays capturable.
void launch_hash_segs(hipStream_t st, const HashSeg* segs, uint32_t nseg, uint32_t blocks,
                      unsigned long long* outs) {
    if (nseg == 0 || blocks == 0) return;
    k_hash_segs<<<blocks, kBlock, 0, st>>>(segs, nseg, outs);
}

void launch_gather_rows(hipStream_t st, const float* src, const uint32_t* idx, uint64_t n,
                        uint32_t base, uint32_t w, float* dst) {
    if (n == 0) return;
    uint64_t blocks = (n * w + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    k_gather_rows<<<(unsigned)blocks, kBlock, 0, st>>>(src, idx, n, base, w, dst);
}

void launch_tri_bounds(hipStream_t st, const float* v, uint64_t n, int32_t* out) {
    if (n == 0) return;
    uint64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 4096) blocks = 4096;
    k_tri_bounds<<<(unsigned)blocks, kBlock, 0, st>>>(v, n, out);
}

void launch_copy_u128(hipStream_t st, const void* src, void* dst, uint64_t n16) {
    if (n16 == 0) return;
    // about one pass of four 16-byte words per lane (a grid-stride walk of
    // few steps moved the most bytes per second)
    const uint64_t blocks = std::min<uint64_t>((n16 + 4ull * kBlock - 1) / (4ull * kBlock), 262144);
    k_copy_u128<<<(unsigned)blocks, kBlock, 0, st>>>(static_cast<const uint4*>(src),
                                                   static_cast<uint4*>(dst), n16);
}

void launch_copy_scattered(hipStream_t st, const void* src, void* dst, uint64_t nblocks) {
    if (nblocks == 0) return;
    k_copy_scattered<<<4096, kBlock, 0, st>>>(static_cast<const uint4*>(src),
                                              static_cast<uint4*>(dst), nblocks);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    if (n == 0) return;
    uint64_t blocks = ((n + 3) / 4 + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    k_fill_u32<<<(unsigned)blocks, kBlock, 0, st>>>(p, n, v);
}
void launch_fill_segs(hipStream_t st, const FillSeg* segs, int nseg) {
    FillSegs F{};
    uint64_t total = 0;
    F.k = 0;
    for (int i = 0; i < nseg && F.k < 4; ++i) {
        if (segs[i].n == 0 || !segs[i].p) continue;
        F.s[F.k++] = segs[i];
        total += (segs[i].n + 3) / 4;
    }
    if (total == 0) return;
    uint64_t blocks = (total + kBlock - 1) / kBlock;
    if (blocks > 65536) blocks = 65536;
    k_fill_segs<<<(unsigned)blocks, kBlock, 0, st>>>(F);
}

// ---------------------------------------------------------------------------
// launchers
namespace {
// Launches over the groups g0 .. g1 - 1 that `use` admits, kGroupSet at a
// time, each group's triangles in its own range of workgroups.
template <class Use, class Launch>
void launch_group_sets(const std::vector<GroupArgs>& groups, Use use, Launch launch) {
    GroupSet S{};
    uint32_t blocks = 0;
    auto flush = [&] {
        if (S.n == 0) return;
        S.start[S.n] = blocks;
        launch(S, blocks);
        S = GroupSet{};
        blocks = 0;
    };
    for (int g = 0; g < (int)groups.size(); ++g) {
        if (!use(groups[g])) continue;
        S.G[S.n] = groups[g];
        S.g[S.n] = g;
        S.start[S.n] = blocks;
        blocks += grid_for(groups[g].ntri, kBlock);
        if (++S.n == kGroupSet) flush();
    }
    flush();
}
}  // namespace

void launch_setup(hipStream_t st, const VoxParams& P, const std::vector<GroupArgs>& groups,
                  const float* verts, float* fnorm, uint32_t* bbox, uint32_t* binCount) {
    launch_group_sets(groups, [](const GroupArgs& G) { return G.ntri > 0; },
                      [&](const GroupSet& S, uint32_t blocks) {
                          k_setup<<<blocks, kBlock, 0, st>>>(P, S, verts, fnorm, bbox, binCount);
                      });
}
void launch_scatter(hipStream_t st, const VoxParams& P, const std::vector<GroupArgs>& groups,
                    const uint32_t* bbox, const uint32_t* binStart, uint32_t* binFill,
                    uint32_t* binTri) {
    launch_group_sets(groups, [](const GroupArgs& G) { return G.ntri > 0 && G.overlaps; },
                      [&](const GroupSet& S, uint32_t blocks) {
                          k_scatter<<<blocks, kBlock, 0, st>>>(P, S, bbox, binStart, binFill,
                                                               binTri);
                      });
}
void launch_activity0(hipStream_t st, const VoxParams& P, int ngroups, const int32_t* grpFill,
                      const uint32_t* binStart, uint32_t* mark) {
    k_activity0<<<grid_for(P.numTiles, kBlock), kBlock, 0, st>>>(P, ngroups, grpFill, binStart, mark);
}
void launch_compact(hipStream_t st, uint64_t n, const uint32_t* mark, const uint32_t* excl,
                    int32_t* tileIdx, int32_t* list, uint32_t* count) {
    if (n == 0) {
        launch_fill_u32(st, count, 1, 0u);
        return;
    }
    k_compact<<<grid_for(n, kBlock), kBlock, 0, st>>>(n, mark, excl, tileIdx, list, count);
}
void launch_mark_children(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                          uint32_t* markDst) {
    launch_mark_children_shifted(st, src, dst, MipShift{}, markDst);
}
void launch_mark_children_shifted(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                                  const MipShift& shift, uint32_t* markDst) {
    if (dst.numTiles == 0) return;
    k_mark_children<<<grid_for(dst.numTiles, kBlock), kBlock, 0, st>>>(src, dst, shift, markDst);
}
void launch_voxelize(hipStream_t st, const VoxParams& P, const GroupArgs& G, int g, int init,
                     uint64_t nActive, const int32_t* list, const float* verts,
                     const float* fnorm, const uint32_t* bbox, const uint32_t* binStart,
                     const uint32_t* binTri, uint32_t* occ, uint32_t* hit,
                     unsigned long long* satCounter, uint4* work, uint32_t* workCount,
                     uint4* rec, uint32_t* recCount, unsigned persistentBlocks) {
    if (nActive == 0) return;
    const unsigned blocks = (unsigned)vox_grid(nActive, persistentBlocks);
    k_voxelize<<<blocks, kBlock, 0, st>>>(P, G, g, init, work, workCount, list, verts, fnorm, bbox,
                                          binStart, binTri, occ, hit, rec, recCount,
                                          rec_seg_cap(nActive, blocks), satCounter);
}
VoxUnitJob voxelize_job(const VoxParams& P, const GroupArgs& G, int init, uint64_t nActive,
                        const float* verts, const float* fnorm, const uint32_t* bbox,
                        const uint32_t* binTri, uint32_t* occ, uint32_t* hit, uint4* work,
                        uint32_t* workCount, uint4* rec, uint32_t* recCount,
                        unsigned persistentBlocks) {
    VoxUnitJob j{};
    j.P = P;
    j.G = G;
    j.init = init;
    j.work = work;
    j.workCount = workCount;
    j.verts = verts;
    j.fnorm = fnorm;
    j.bbox = bbox;
    j.binTri = binTri;
    j.occ = occ;
    j.hit = hit;
    j.rec = rec;
    j.recCount = recCount;
    // the job's record segments are sized for its own launch shape; units
    // go to segment (unit % kRecSegs), which that sizing always holds
    j.segCap = rec_seg_cap(nActive, vox_grid(nActive, persistentBlocks));
    return j;
}
void launch_voxelize_batch(hipStream_t st, const VoxUnitJob* jobs, uint32_t njobs,
                           uint32_t* unitStart, uint32_t* queue, unsigned blocks) {
    if (njobs == 0 || blocks == 0) return;
    k_voxelize_batch_units<<<1, kBlock, 0, st>>>(jobs, njobs, unitStart, queue);
    k_voxelize_batch<<<blocks, kBlock, 0, st>>>(jobs, njobs, unitStart, queue);
}
void launch_group_tiles(hipStream_t st, const VoxParams& P, int g, uint64_t nActive,
                        const int32_t* list, const uint32_t* binStart, uint4* work,
                        uint32_t* workCount, uint32_t* recCount) {
    const FillSeg z[2] = {{workCount, 1, 0u}, {recCount, kRecSegs, 0u}};
    launch_fill_segs(st, z, 2);
    if (nActive == 0) return;
    k_group_tiles<<<grid_for(nActive, kBlock), kBlock, 0, st>>>(P, g, nActive, list, binStart,
                                                                 work, workCount);
}
void launch_colour(hipStream_t st, const VoxParams& P, const GroupArgs& G, int init,
                   uint64_t nActive, const uint4* rec, const uint32_t* recCount,
                   const int32_t* list, const float* verts, const float* fnorm, uint32_t* col,
                   uint32_t* qn, float* nrm, const uint32_t* fillw, unsigned persistentBlocks) {
    const uint64_t vblocks = vox_grid(nActive, persistentBlocks);
    k_colour<<<persistentBlocks, kBlock, 0, st>>>(P, G, init, rec, recCount,
                                                  rec_seg_cap(nActive, vblocks), list, verts,
                                                  fnorm, col, qn, nrm, fillw);
}
void launch_fill(hipStream_t st, const VoxParams& P, int g, int init, uint64_t nActive,
                 const int32_t* list, const int32_t* tileIdx, const uint32_t* binStart,
                 uint32_t* col, uint32_t* qn, float* nrm, const uint32_t* hit, uint32_t* fillw) {
    if (nActive == 0) return;
    k_fill<<<(unsigned)((nActive + kFillWaves - 1) / kFillWaves), kBlock, 0, st>>>(
        P, g, init, nActive, list, tileIdx, binStart, col, qn, nrm, hit, fillw);
}
void launch_mip(hipStream_t st, const LevelMap& src, const LevelMap& dst, uint64_t nActive,
                const int32_t* listDst, const uint32_t* colSrc, const float* nrmSrc,
                uint32_t* colDst, uint32_t* qnDst, float* nrmDst) {
    launch_mip_shifted(st, src, dst, MipShift{}, nActive, listDst, colSrc, nrmSrc, colDst, qnDst,
                       nrmDst);
}
void launch_mip_shifted(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                        const MipShift& shift, uint64_t nActive, const int32_t* listDst,
                        const uint32_t* colSrc, const float* nrmSrc, uint32_t* colDst,
                        uint32_t* qnDst, float* nrmDst) {
    if (nActive == 0) return;
    k_mip<<<(unsigned)nActive, kBlock, 0, st>>>(src, dst, shift, listDst, colSrc, nrmSrc, colDst,
                                                qnDst, nrmDst);
}
bool mip_shift_for(const LevelMap& src, const LevelMap& dst, const ChunkGeom& cg, int ks, int kd,
                   MipShift& out) {
    out = MipShift{};
    for (int a = 0; a < 3; ++a) {
        const uint32_t Cd = cg.C[a] >> kd, Cs = cg.C[a] >> ks;
        if (Cd == 0 || Cd % kTile != 0 || Cs < 2 * Cd) return false;
        const uint32_t off = Cs - 2 * Cd;
        if (off % (2 * kTile) != 0) return false;
        if (dst.dims[a] == 0) return false;
        // the last voxel's second source (the largest) below the source's end
        const uint64_t X = dst.dims[a] - 1, i = X / Cd, x = X - i * Cd;
        if (i * Cs + 2 * x + 1 >= src.dims[a]) return false;
        out.cd[a] = Cd;
        out.off[a] = off;
    }
    return true;
}
void launch_node_sample_all(hipStream_t st, const SampleDepths& S) {
    if (S.L <= 0 || S.start[S.L] == 0) return;
    k_node_sample_all<<<grid_for(S.start[S.L], kBlock), kBlock, 0, st>>>(S);
}
void launch_node_compare(hipStream_t st, const LevelMap& lm, uint64_t nActive, const int32_t* list,
                         const uint32_t* col, uint32_t Rd, const uint32_t B[3],
                         const uint32_t border[3], uint32_t xLimit, uint32_t* cc, uint32_t* flag) {
    if (nActive == 0) return;
    k_node_compare<<<(unsigned)((nActive + kNcTiles - 1) / kNcTiles), kBlock, 0, st>>>(
        lm, list, nActive, col, Rd, B[0], B[1], B[2], border[0], border[1], border[2], xLimit, cc,
        flag);
}
void launch_range_counts(hipStream_t st, const uint32_t* excl, const uint64_t* offs, int n,
                         uint32_t* out) {
    if (n <= 0) return;
    RangeOffs o{};
    for (int d = 0; d <= n && d <= kMaxLevels; ++d) o.v[d] = offs[d];
    k_range_counts<<<1, 64, 0, st>>>(excl, o, n < kMaxLevels ? n : kMaxLevels, out);
}
void launch_brick_assign_all(hipStream_t st, const uint32_t* flag, const uint32_t* excl,
                             const DepthOut& D) {
    if (D.L <= 0) return;
    const uint64_t all = D.off[D.L];
    if (all == 0) return;
    k_brick_assign_all<<<grid_for(all, kBlock), kBlock, 0, st>>>(all, flag, excl, D);
}
void launch_brick_gather_all(hipStream_t st, const GatherDepths& G) {
    if (G.n <= 0 || G.start[G.n] == 0) return;
    k_brick_gather_all<<<(unsigned)G.start[G.n], kBlock, 0, st>>>(G);
}
void launch_collapse(hipStream_t st, uint32_t Rd, int childIsLeaf, uint8_t* types,
                     const uint8_t* childTypes) {
    const uint64_t nn = (uint64_t)Rd * Rd * Rd;
    k_collapse<<<grid_for(nn, kBlock), kBlock, 0, st>>>(Rd, childIsLeaf, types, childTypes);
}
void launch_occupancy(hipStream_t st, const VoxParams& P, uint64_t nActive, const int32_t* list,
                      const uint32_t* occ, uint32_t* bitmap) {
    if (nActive == 0) return;
    k_occupancy<<<(unsigned)nActive, kBlock, 0, st>>>(P, list, occ, bitmap);
}
bool launch_occupancy_rows(hipStream_t st, const VoxParams& P, const int32_t* tileIdx,
                           const uint32_t* occ, uint32_t* bitmap) {
    if ((P.S[0] & 31u) != 0u || P.TG[0] * 4u > 65536u) return false;
    const uint64_t rows = (uint64_t)P.TG[1] * P.TG[2];
    if (rows == 0) return true;
    if ((P.S[0] & 127u) == 0u)
        k_occupancy_quads<<<(unsigned)rows, kOccQuadBlock, P.TG[0] * 4u, st>>>(P, tileIdx, occ,
                                                                              bitmap);
    else
        k_occupancy_rows<<<(unsigned)rows, kBlock, P.TG[0] * 4u, st>>>(P, tileIdx, occ, bitmap);
    return true;
}
void launch_dense_level(hipStream_t st, const LevelMap& lm, const uint32_t* col,
                        const float* nrm, uint32_t* outCol, float* outNrm) {
    const uint64_t n = (uint64_t)lm.dims[0] * lm.dims[1] * lm.dims[2];
    if (n == 0) return;
    k_dense_level<<<grid_for(n, kBlock), kBlock, 0, st>>>(lm, col, nrm, outCol, outNrm);
}
uint64_t scan_tmp_words(uint64_t n) { return 2 * ((n + kScanTile - 1) / kScanTile + 1); }
bool ensure_scan_tmp(DevBuf& b, uint64_t n, hipStream_t st, std::stringstream& err,
                     const char* what) {
    const size_t cap = b.cap;
    if (!b.ensure(scan_tmp_words(n) * 4, err, what)) return false;
    if (b.cap != cap) launch_fill_u32(st, b.as<uint32_t>(), b.cap / 4, 0u);
    return true;
}
static void scan_launch(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n,
                        uint32_t* tmp, uint32_t* clearIn, uint32_t* err) {
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) {
        launch_fill_u32(st, out, 1, 0u);
        return;
    }
    k_scan<<<(unsigned)nb, kBlock, 0, st>>>(in, n, (uint32_t)nb, out, clearIn,
                                             reinterpret_cast<unsigned long long*>(tmp), err);
}
void launch_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                 uint32_t* err) {
    scan_launch(st, in, out, n, tmp, nullptr, err);
}
void launch_scan_clear(hipStream_t st, uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                       uint32_t* err) {
    scan_launch(st, in, out, n, tmp, in, err);
}


void launch_mip_source_count(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                             uint64_t nActive, const int32_t* listDst, unsigned long long* out) {
    if (nActive == 0) return;
    k_mip_source_count<<<grid_for(nActive, kBlock), kBlock, 0, st>>>(src, dst, nActive, listDst,
                                                                    out);
}

void launch_chunk_mark(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                       const ChunkGeom& cg, int ks, int kd, uint32_t* markDst) {
    if (dst.numTiles == 0) return;
    k_chunk_mark<<<grid_for(dst.numTiles, kBlock), kBlock, 0, st>>>(src, dst, cg, ks, kd, markDst);
}
void launch_chunk_mip(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                      const ChunkGeom& cg, int ks, int kd, uint64_t nActive,
                      const int32_t* listDst, const uint32_t* colSrc, const float* nrmSrc,
                      uint32_t* colDst, uint32_t* qnDst, float* nrmDst) {
    if (nActive == 0) return;
    k_chunk_mip<<<(unsigned)nActive, kBlock, 0, st>>>(src, dst, cg, ks, kd, listDst, colSrc, nrmSrc,
                                                      colDst, qnDst, nrmDst);
}
void launch_chunk_node_type(hipStream_t st, const SampleDepths& S, const ChunkGeom& cg) {
    const uint64_t nn = S.start[S.L];
    if (nn == 0) return;
    const unsigned grid = (unsigned)std::min<uint64_t>((nn + kChunkWaves - 1) / kChunkWaves, 1u << 20);
    k_chunk_node_type<<<grid, kBlock, 0, st>>>(S, cg);
}
void launch_chunk_brick_keys(hipStream_t st, const uint32_t* flag, const ChunkGeom& cg, int k,
                             uint32_t Rd, const uint32_t B[3], const uint32_t border[3],
                             unsigned long long* keys, uint32_t* nodes, uint32_t* count) {
    const uint64_t nn = (uint64_t)Rd * Rd * Rd;
    k_chunk_brick_keys<<<grid_for(nn, kBlock), kBlock, 0, st>>>(flag, cg, k, Rd, B[0], B[1], B[2],
                                                                border[0], border[1], border[2],
                                                                keys, nodes, count);
}
void launch_chunk_brick_assign(hipStream_t st, uint64_t nodes, const uint32_t* flag,
                               const uint32_t* sorted, uint32_t nb, uint8_t* types,
                               uint32_t* bidx, uint32_t* brickNode) {
    if (nodes) k_chunk_types<<<grid_for(nodes, kBlock), kBlock, 0, st>>>(nodes, flag, types, bidx);
    if (nb) k_chunk_bidx<<<grid_for(nb, kBlock), kBlock, 0, st>>>(nb, sorted, bidx, brickNode);
}

}  // namespace voxmi
