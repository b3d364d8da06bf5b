// vox_tiler.hip -- assignment of triangles to the 2-D grid of octree tiles
// (C ABI in include/voxmi_tiler.h).
//
// TriangleOctTreeGrid::assignTriangles (main.cu 3986-4030) gives each
// triangle the tiles between the smallest and largest cell its vertices
// reach, where a vertex within the voxel border of a cell edge also reaches
// the neighbouring cell (computeMinMaxGridCell, main.cu 4078-4107) and a
// vertex left of / below the grid origin reaches cell 0.  Here one thread per
// triangle computes that range; per-tile counts and z extents (GridTree::add,
// main.cu 3896-3915) use wave-aggregated atomics; every triangle then writes
// its (tile, triangle) pairs at its own offset, so the pairs lie in triangle
// order, and a stable radix sort by tile leaves each tile's triangles in
// input order, the order the reference appends them in.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

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

// Counts per tile and per triangle, and the tiles' z extents.
__global__ void k_tile_count(const float* __restrict__ verts, uint64_t ntri, Grid g,
                             uint32_t* __restrict__ cnt, uint32_t* __restrict__ triCnt,
                             int32_t* __restrict__ zmin, int32_t* __restrict__ zmax) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const float* v = verts + 9 * t;
    uint32_t x0, x1, y0, y1;
    tri_range(v, g, x0, x1, y0, y1);
    triCnt[t] = (x1 - x0 + 1u) * (y1 - y0 + 1u);
    const int32_t zlo = ord(fminf(fminf(v[2], v[5]), v[8]));
    const int32_t zhi = ord(fmaxf(fmaxf(v[2], v[5]), v[8]));
    const int lane = (int)(threadIdx.x & 63u);
    uint32_t x = x0, y = y0;
    bool act = true;
    while (__ballot(act)) {
        const uint64_t key = (uint64_t)y * g.dim[0] + x;
        uint64_t pend = __ballot(act);
        while (pend) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint64_t lk = shfl64(key, l);
            const uint64_t same = __ballot(act && key == lk);
            if (lane == l) atomicAdd(&cnt[lk], (uint32_t)__popcll(same));
            if (act && key == lk) {
                atomicMin(&zmin[lk], zlo);
                atomicMax(&zmax[lk], zhi);
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

// The (tile, triangle) pairs of triangle t at its offset: pairs in
// triangle order, tiles row by row within a triangle.
__global__ void k_tile_pairs(const float* __restrict__ verts, uint64_t ntri, Grid g,
                             const uint32_t* __restrict__ triOff, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    uint32_t x0, x1, y0, y1;
    tri_range(verts + 9 * t, g, x0, x1, y0, y1);
    uint32_t o = triOff[t];
    for (uint32_t y = y0; y <= y1; ++y)
        for (uint32_t x = x0; x <= x1; ++x) {
            keys[o] = y * g.dim[0] + x;
            vals[o] = (uint32_t)t;
            ++o;
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

// For every tile and group boundary k: the position in the tile's ordered
// list of its first triangle with index >= first[k].
__global__ void k_group_bounds(const uint32_t* __restrict__ start, const uint32_t* __restrict__ tris,
                               uint64_t ntiles, const uint32_t* __restrict__ first, uint32_t nb,
                               uint32_t* __restrict__ out) {
    const uint64_t total = ntiles * nb;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t tile = e / nb;
        const uint32_t v = first[e - tile * nb];
        uint32_t lo = start[tile], hi = start[tile + 1];
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (tris[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        out[e] = lo;
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
    DevBuf triCnt, triOff, triScan, keys, keys2, vals, sortTmp;
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
                   &t->scanTmp, &t->hostVerts, &t->triCnt, &t->triOff, &t->triScan, &t->keys,
                   &t->keys2, &t->vals, &t->sortTmp};
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
    if (n >= 0xffffffffull) return t->fail("tiler: too many tiles"), 0;
    t->ntiles = n;
    hipStream_t st = t->st;
    const uint64_t nt = ntri ? ntri : 1;
    if (!t->cnt.ensure(4 * (n + 1), t->err, "tile counts") ||
        !t->start.ensure(4 * (n + 1), t->err, "tile offsets") ||
        !t->zmin.ensure(4 * n, t->err, "z min") || !t->zmax.ensure(4 * n, t->err, "z max") ||
        !t->zminf.ensure(4 * n, t->err, "z min") || !t->zmaxf.ensure(4 * n, t->err, "z max") ||
        !ensure_scan_tmp(t->scanTmp, n, st, t->err, "scan") ||
        !t->triCnt.ensure(4 * (nt + 1), t->err, "triangle counts") ||
        !t->triOff.ensure(4 * (nt + 1), t->err, "triangle offsets") ||
        !ensure_scan_tmp(t->triScan, nt, st, t->err, "scan"))
        return 0;
    launch_fill_u32(st, t->cnt.as<uint32_t>(), n, 0u);
    k_fill_i32<<<grid_for(n), kB, 0, st>>>(t->zmin.as<int32_t>(), n, 0x7fffffff);
    k_fill_i32<<<grid_for(n), kB, 0, st>>>(t->zmax.as<int32_t>(), n, (int32_t)0x80000000);
    if (ntri) {
        k_tile_count<<<(unsigned)((ntri + kB - 1) / kB), kB, 0, st>>>(
            verts, ntri, g, t->cnt.as<uint32_t>(), t->triCnt.as<uint32_t>(),
            t->zmin.as<int32_t>(), t->zmax.as<int32_t>());
        launch_scan(st, t->triCnt.as<uint32_t>(), t->triOff.as<uint32_t>(), ntri,
                    t->triScan.as<uint32_t>());
    }
    launch_scan(st, t->cnt.as<uint32_t>(), t->start.as<uint32_t>(), n, t->scanTmp.as<uint32_t>());
    uint32_t total = 0;
    if (!t->check(hipMemcpyAsync(&total, t->start.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, st),
                  "pairs") ||
        !t->check(hipStreamSynchronize(st), "synchronize"))
        return 0;
    t->pairs = total;
    if (!t->tris.ensure(4ull * total + 4, t->err, "tile triangles")) return 0;
    if (total) {
        if (!t->keys.ensure(4ull * total, t->err, "pair keys") ||
            !t->keys2.ensure(4ull * total, t->err, "pair keys") ||
            !t->vals.ensure(4ull * total, t->err, "pair triangles"))
            return 0;
        k_tile_pairs<<<(unsigned)((ntri + kB - 1) / kB), kB, 0, st>>>(
            verts, ntri, g, t->triOff.as<uint32_t>(), t->keys.as<uint32_t>(),
            t->vals.as<uint32_t>());
        unsigned bits = 1;
        while (bits < 32 && (1ull << bits) < n) ++bits;
        size_t tmpBytes = 0;
        if (!t->check(rocprim::radix_sort_pairs(nullptr, tmpBytes, t->keys.as<uint32_t>(),
                                                t->keys2.as<uint32_t>(), t->vals.as<uint32_t>(),
                                                t->tris.as<uint32_t>(), (size_t)total, 0u, bits,
                                                st),
                      "sort size") ||
            !t->sortTmp.ensure(tmpBytes, t->err, "sort storage") ||
            !t->check(rocprim::radix_sort_pairs(t->sortTmp.p, tmpBytes, t->keys.as<uint32_t>(),
                                                t->keys2.as<uint32_t>(), t->vals.as<uint32_t>(),
                                                t->tris.as<uint32_t>(), (size_t)total, 0u, bits,
                                                st),
                      "sort"))
            return 0;
    }
    k_unord<<<grid_for(n), kB, 0, st>>>(t->zmin.as<int32_t>(), n, t->zminf.as<float>());
    k_unord<<<grid_for(n), kB, 0, st>>>(t->zmax.as<int32_t>(), n, t->zmaxf.as<float>());
    return t->check(hipGetLastError(), "tiler kernels") &&
           t->check(hipStreamSynchronize(st), "synchronize");
}

int vox_tiler_group_bounds(vox_tiler* t, const uint32_t* first, uint32_t nbounds, uint32_t* out) {
    if (!t || !t->valid || !first || !out || nbounds == 0) return 0;
    if (t->ntiles == 0) return t->fail("tiler: assign first"), 0;
    const uint64_t total = t->ntiles * nbounds;
    DevBuf f, o;
    bool ok = f.ensure(4ull * nbounds, t->err, "bounds") && o.ensure(4 * total, t->err, "bounds") &&
              t->check(hipMemcpyAsync(f.p, first, 4ull * nbounds, hipMemcpyHostToDevice, t->st),
                       "bounds upload");
    if (ok) {
        k_group_bounds<<<grid_for(total), kB, 0, t->st>>>(t->start.as<uint32_t>(),
                                                          t->tris.as<uint32_t>(), t->ntiles,
                                                          f.as<uint32_t>(), nbounds,
                                                          o.as<uint32_t>());
        ok = t->check(hipGetLastError(), "bounds kernel") &&
             t->check(hipMemcpyAsync(out, o.p, 4 * total, hipMemcpyDeviceToHost, t->st),
                      "bounds") &&
             t->check(hipStreamSynchronize(t->st), "synchronize");
    }
    f.release();
    o.release();
    return ok ? 1 : 0;
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
