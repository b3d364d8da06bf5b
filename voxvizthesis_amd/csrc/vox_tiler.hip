// vox_tiler.hip -- assignment of triangles to the 2-D grid of octree tiles
// (C ABI in include/voxmi_tiler.h).
//
// TriangleOctTreeGrid::assignTriangles (main.cu 3986-4030) gives each
// triangle the tiles between the smallest and largest cell its vertices
// reach, where a vertex within the voxel border of a cell edge also reaches
// the neighbouring cell (computeMinMaxGridCell, main.cu 4078-4107) and a
// vertex left of / below the grid origin reaches cell 0.  Here one thread per
// triangle computes that range, the per-tile counts and placements use
// wave-aggregated atomics over a prefix sum, and each tile also gets the z
// extent of its triangles' vertices (GridTree::add, main.cu 3896-3915).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/voxmi_tiler.h"
#include "vox_internal.h"

using namespace voxmi;

namespace {

constexpr int kB = 256;

struct Grid {
    float origin[2], size[2], border[2];
    uint32_t dim[2];
};

// computeMinMaxGridCell for one coordinate.
__device__ __forceinline__ void cell_range(float v, float origin, float size, float border,
                                           uint32_t dim, uint32_t& mn, uint32_t& mx) {
    if (v < origin) {
        mn = mx = 0u;
        return;
    }
    const float dist = v - origin;
    const float q = dist / size;
    mn = mx = q < 4294967040.0f ? (uint32_t)q : 0xffffff00u;
    const float maxBorder = (float)(mn + 1u) * size;
    const float d = maxBorder - dist;
    if (d <= border) {
        mx = mn + 1u;
    } else if (mn != 0u) {
        const float minBorder = maxBorder - size;
        if (dist - minBorder <= border) {
            mx = mn;
            mn = mn - 1u;
        }
    }
    if (mn >= dim) mn = dim - 1u;
    if (mx >= dim) mx = dim - 1u;
}

__device__ __forceinline__ void tri_range(const float* __restrict__ v, const Grid& g,
                                          uint32_t& x0, uint32_t& x1, uint32_t& y0, uint32_t& y1) {
    x0 = y0 = 0xffffffffu;
    x1 = y1 = 0u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t a, b, c, d;
        cell_range(v[3 * k], g.origin[0], g.size[0], g.border[0], g.dim[0], a, b);
        cell_range(v[3 * k + 1], g.origin[1], g.size[1], g.border[1], g.dim[1], c, d);
        x0 = a < x0 ? a : x0;
        x1 = b > x1 ? b : x1;
        y0 = c < y0 ? c : y0;
        y1 = d > y1 ? d : y1;
    }
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int l) {
    const uint32_t lo = __shfl((uint32_t)v, l), hi = __shfl((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Order-preserving int image of a float (for atomic min / max).
__device__ __forceinline__ int32_t ord(float f) {
    const int32_t i = __float_as_int(f);
    return i >= 0 ? i : (int32_t)(0x80000000u - (uint32_t)i) - 1;
}

// pass 0: counts and z extents; pass 1: placement.
template <int PASS>
__global__ void k_tile_assign(const float* __restrict__ verts, uint64_t ntri, Grid g,
                              uint32_t* __restrict__ cnt, const uint32_t* __restrict__ start,
                              uint32_t* __restrict__ tris, int32_t* __restrict__ zmin,
                              int32_t* __restrict__ zmax) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const float* v = verts + 9 * t;
    uint32_t x0, x1, y0, y1;
    tri_range(v, g, x0, x1, y0, y1);
    int32_t zlo = 0, zhi = 0;
    if (PASS == 0) {
        zlo = ord(fminf(fminf(v[2], v[5]), v[8]));
        zhi = ord(fmaxf(fmaxf(v[2], v[5]), v[8]));
    }
    const int lane = (int)(threadIdx.x & 63u);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t x = x0, y = y0;
    bool act = true;
    while (__ballot(act)) {
        const uint64_t key = (uint64_t)y * g.dim[0] + x;
        uint64_t pend = __ballot(act);
        while (pend) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint64_t lk = shfl64(key, l);
            const uint64_t same = __ballot(act && key == lk);
            if (PASS == 0) {
                if (lane == l) atomicAdd(&cnt[lk], (uint32_t)__popcll(same));
                if (act && key == lk) {
                    atomicMin(&zmin[lk], zlo);
                    atomicMax(&zmax[lk], zhi);
                }
            } else {
                uint32_t base = 0;
                if (lane == l) base = atomicAdd(&cnt[lk], (uint32_t)__popcll(same));
                base = __shfl(base, l);
                if (act && key == lk)
                    tris[start[lk] + base + (uint32_t)__popcll(same & below)] = (uint32_t)t;
            }
            pend &= ~same;
        }
        if (act) {
            if (++x > x1) {
                x = x0;
                if (++y > y1) act = false;
            }
        }
    }
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_unord(const int32_t* __restrict__ in, uint64_t n, float* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t o = in[i];
        const int32_t b = o >= 0 ? o : (int32_t)(0x80000000u - (uint32_t)(o + 1));
        out[i] = __int_as_float(b);
    }
}

unsigned grid_for(uint64_t n) {
    const uint64_t g = (n + kB - 1) / kB;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace

struct vox_tiler {
    bool valid = false;
    hipStream_t st = nullptr;
    uint64_t ntiles = 0, pairs = 0;
    DevBuf cnt, start, tris, zmin, zmax, zminf, zmaxf, scanTmp, hostVerts;
    std::stringstream err;
    std::string errCache;
    bool fail(const std::string& m) {
        err << m << std::endl;
        return false;
    }
    bool check(hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        err << what << " failed: " << hipGetErrorString(e) << std::endl;
        return false;
    }
};

extern "C" {

vox_tiler* vox_tiler_create(void) {
    vox_tiler* t = new (std::nothrow) vox_tiler;
    if (!t) return nullptr;
    if (!t->check(hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking), "hipStreamCreate"))
        return t;
    t->valid = true;
    return t;
}

void vox_tiler_destroy(vox_tiler* t) {
    if (!t) return;
    if (t->st) (void)hipStreamSynchronize(t->st);
    DevBuf* b[] = {&t->cnt, &t->start, &t->tris, &t->zmin, &t->zmax, &t->zminf, &t->zmaxf,
                   &t->scanTmp, &t->hostVerts};
    for (DevBuf* x : b) x->release();
    if (t->st) (void)hipStreamDestroy(t->st);
    delete t;
}

const char* vox_tiler_get_error(vox_tiler* t) {
    if (!t) return "null tiler";
    t->errCache = t->err.str();
    return t->errCache.c_str();
}

int vox_tiler_assign(vox_tiler* t, const float* verts, uint64_t ntri, int32_t device,
                     const vox_tile_grid* grid) {
    if (!t || !t->valid || !grid) return 0;
    if (ntri && !verts) return t->fail("tiler: no vertices"), 0;
    if (ntri && !device) {
        if (!t->hostVerts.ensure(36 * ntri, t->err, "vertices") ||
            !t->check(hipMemcpyAsync(t->hostVerts.p, verts, 36 * ntri, hipMemcpyHostToDevice, t->st),
                      "vertex upload"))
            return 0;
        verts = t->hostVerts.as<float>();
    }
    Grid g;
    for (int a = 0; a < 2; ++a) {
        g.origin[a] = grid->origin[a];
        g.size[a] = grid->size[a];
        g.border[a] = grid->border[a];
        g.dim[a] = grid->dim[a];
        if (!(g.size[a] > 0.0f) || g.dim[a] == 0 || g.dim[a] > 65536)
            return t->fail("tiler: bad grid"), 0;
    }
    const uint64_t n = (uint64_t)g.dim[0] * g.dim[1];
    t->ntiles = n;
    hipStream_t st = t->st;
    if (!t->cnt.ensure(4 * (n + 1), t->err, "tile counts") ||
        !t->start.ensure(4 * (n + 1), t->err, "tile offsets") ||
        !t->zmin.ensure(4 * n, t->err, "z min") || !t->zmax.ensure(4 * n, t->err, "z max") ||
        !t->zminf.ensure(4 * n, t->err, "z min") || !t->zmaxf.ensure(4 * n, t->err, "z max") ||
        !t->scanTmp.ensure(4 * scan_tmp_words(n), t->err, "scan"))
        return 0;
    launch_fill_u32(st, t->cnt.as<uint32_t>(), n, 0u);
    k_fill_i32<<<grid_for(n), kB, 0, st>>>(t->zmin.as<int32_t>(), n, 0x7fffffff);
    k_fill_i32<<<grid_for(n), kB, 0, st>>>(t->zmax.as<int32_t>(), n, (int32_t)0x80000000);
    if (ntri)
        k_tile_assign<0><<<(unsigned)((ntri + kB - 1) / kB), kB, 0, st>>>(
            verts, ntri, g, t->cnt.as<uint32_t>(), nullptr, nullptr, t->zmin.as<int32_t>(),
            t->zmax.as<int32_t>());
    launch_scan(st, t->cnt.as<uint32_t>(), t->start.as<uint32_t>(), n, t->scanTmp.as<uint32_t>());
    uint32_t total = 0;
    if (!t->check(hipMemcpyAsync(&total, t->start.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, st),
                  "pairs") ||
        !t->check(hipStreamSynchronize(st), "synchronize"))
        return 0;
    t->pairs = total;
    if (!t->tris.ensure(4ull * total + 4, t->err, "tile triangles")) return 0;
    launch_fill_u32(st, t->cnt.as<uint32_t>(), n, 0u);
    if (ntri)
        k_tile_assign<1><<<(unsigned)((ntri + kB - 1) / kB), kB, 0, st>>>(
            verts, ntri, g, t->cnt.as<uint32_t>(), t->start.as<uint32_t>(), t->tris.as<uint32_t>(),
            nullptr, nullptr);
    k_unord<<<grid_for(n), kB, 0, st>>>(t->zmin.as<int32_t>(), n, t->zminf.as<float>());
    k_unord<<<grid_for(n), kB, 0, st>>>(t->zmax.as<int32_t>(), n, t->zmaxf.as<float>());
    return t->check(hipGetLastError(), "tiler kernels") &&
           t->check(hipStreamSynchronize(st), "synchronize");
}

int vox_tiler_get(vox_tiler* t, vox_tile_lists* out) {
    if (!t || !t->valid || !out) return 0;
    out->ntiles = t->ntiles;
    out->pairs = t->pairs;
    out->tile_start = t->start.as<uint32_t>();
    out->tile_tris = t->tris.as<uint32_t>();
    out->tile_zmin = t->zminf.as<float>();
    out->tile_zmax = t->zmaxf.as<float>();
    return 1;
}

}  // extern "C"
