// vox_terrain.hip -- procedural terrain on the device (C ABI in
// include/voxmi_terrain.h): the seeded heightfield of the paged workloads,
// its per-tile triangle soups and height ranges, without a host copy of the
// whole mesh.
//
// A 32769^2 heightfield is 2.1 G triangles (77 GB as a soup), so the
// large-scale served paging builds each tile's triangles from a resident
// height grid when the pager asks for the tile.  The arithmetic is the
// scene generator's (voxvizthesis_amd/scenes.py: fBm of 2-D gradient noise
// with a quintic fade in float64, heights rounded to float32; vertex normals
// by central differences in float32, one-sided at the grid's rim; the UV of
// a triangle at the texel centre of its mean height on a 1-row ramp), in the
// same operation order, so a tile equals the host generator's mesh bit for
// bit on the same grid.  A ridged term (sum of (1 - |noise|)^2 over octaves)
// can be added to the fBm.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/voxmi_terrain.h"
#include "vox_internal.h"

#pragma clang fp contract(off)

using namespace voxmi;

namespace {

thread_local std::string t_err;

bool fail(const std::string& m) {
    t_err = m;
    return false;
}
bool check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    return fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Fisher-Yates shuffle of 0..255 driven by splitmix64(seed), doubled
// (Steele, Lea, Flood 2014 for the generator).
void permutation(uint64_t seed, uint8_t* p512) {
    int32_t p[256];
    for (int i = 0; i < 256; ++i) p[i] = i;
    uint64_t s = seed;
    for (int i = 255; i > 0; --i) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const int j = (int)(z % (uint64_t)(i + 1));
        const int32_t t = p[i];
        p[i] = p[j];
        p[j] = t;
    }
    for (int i = 0; i < 512; ++i) p512[i] = (uint8_t)p[i & 255];
}

// passed by value as a kernel argument (1 KB, well inside the kernel
// argument segment)
struct Perms {
    uint8_t fbm[512];
    uint8_t ridged[512];
};

__device__ __forceinline__ double grad(const uint8_t* __restrict__ perm, int64_t ix, int64_t iy,
                                       double dx, double dy) {
    // gradients (1,1) (-1,1) (1,-1) (-1,-1) (1,0) (-1,0) (0,1) (0,-1)
    const int h = perm[perm[ix] + iy] & 7;
    const double gx = h == 0 || h == 2 || h == 4 ? 1.0 : (h == 1 || h == 3 || h == 5 ? -1.0 : 0.0);
    const double gy = h == 0 || h == 1 || h == 6 ? 1.0 : (h == 2 || h == 3 || h == 7 ? -1.0 : 0.0);
    return gx * dx + gy * dy;
}
__device__ __forceinline__ double fade(double t) {
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0);
}
__device__ double gradient_noise(const uint8_t* __restrict__ perm, double x, double y) {
    const double fx = floor(x), fy = floor(y);
    int64_t xi = (int64_t)fx, yi = (int64_t)fy;
    const double xf = x - (double)xi, yf = y - (double)yi;
    xi &= 255;
    yi &= 255;
    const double n00 = grad(perm, xi, yi, xf, yf);
    const double n10 = grad(perm, xi + 1, yi, xf - 1.0, yf);
    const double n01 = grad(perm, xi, yi + 1, xf, yf - 1.0);
    const double n11 = grad(perm, xi + 1, yi + 1, xf - 1.0, yf - 1.0);
    const double u = fade(xf), v = fade(yf);
    const double nx0 = n00 + u * (n10 - n00);
    const double nx1 = n01 + u * (n11 - n01);
    return nx0 + v * (nx1 - nx0);
}

// 0.5 E + 0.2 E fbm(x/E, y/E) [+ r E ridged(x/E, y/E)], 6 octaves,
// lacunarity 2, gain 0.5, base frequency 4.
__global__ void k_heights(float* __restrict__ out, uint32_t n1, double extent, double step,
                          double ridgedAmp, Perms P) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint64_t)n1 * n1) return;
    const uint32_t j = (uint32_t)(k / n1), i = (uint32_t)(k - (uint64_t)j * n1);
    const double x = ((double)i * step) / extent, y = ((double)j * step) / extent;
    double total = 0.0, amp = 1.0, freq = 4.0, norm = 0.0;
    double rt = 0.0;
    for (int o = 0; o < 6; ++o) {
        total += amp * gradient_noise(P.fbm, x * freq, y * freq);
        if (ridgedAmp != 0.0) {
            const double r = 1.0 - fabs(gradient_noise(P.ridged, x * freq, y * freq));
            rt += amp * (r * r);
        }
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    double h = 0.5 * extent + (0.2 * extent) * (total / norm);
    if (ridgedAmp != 0.0) h = h + (ridgedAmp * extent) * (rt / norm);
    out[k] = (float)h;
}

// Triangles of cells [x0, x0 + cw) x [y0, y0 + ch), row by row, two per
// cell: (v00, v10, v11), (v00, v11, v01).
__global__ void k_tile(const float* __restrict__ H, uint32_t n1, double step, float dxf,
                       uint32_t x0, uint32_t y0, uint32_t cw, uint32_t ch, double lo, double hi,
                       uint32_t texW, float* __restrict__ verts, float* __restrict__ vn,
                       float* __restrict__ uv) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (uint64_t)cw * ch) return;
    const uint32_t cj = (uint32_t)(c / cw), ci = (uint32_t)(c - (uint64_t)cj * cw);
    const uint32_t i = x0 + ci, j = y0 + cj;
    // corners (x, y): 00 = (i, j), 10 = (i+1, j), 11 = (i+1, j+1), 01 = (i, j+1)
    const uint32_t ix[4] = {i, i + 1, i + 1, i}, iy[4] = {j, j, j + 1, j + 1};
    float p[4][3], nrm[4][3];
    const float two = 2.0f;
    for (int q = 0; q < 4; ++q) {
        const uint32_t a = ix[q], b = iy[q];
        p[q][0] = (float)((double)a * step);
        p[q][1] = (float)((double)b * step);
        p[q][2] = H[(uint64_t)b * n1 + a];
        // central differences, the rim padded with the edge value
        const uint32_t am = a > 0 ? a - 1 : 0, ap = a + 1 < n1 ? a + 1 : n1 - 1;
        const uint32_t bm = b > 0 ? b - 1 : 0, bp = b + 1 < n1 ? b + 1 : n1 - 1;
        const float dhdx = (H[(uint64_t)b * n1 + ap] - H[(uint64_t)b * n1 + am]) / (two * dxf);
        const float dhdy = (H[(uint64_t)bp * n1 + a] - H[(uint64_t)bm * n1 + a]) / (two * dxf);
        const float nx = -dhdx, ny = -dhdy, nz = 1.0f;
        const float l = sqrtf((nx * nx + ny * ny) + nz * nz);
        nrm[q][0] = nx / l;
        nrm[q][1] = ny / l;
        nrm[q][2] = nz / l;
    }
    const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
    for (int t = 0; t < 2; ++t) {
        const uint64_t o = 2 * c + t;
        double hs = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int q = tri[t][k];
            for (int a = 0; a < 3; ++a) {
                verts[9 * o + 3 * k + a] = p[q][a];
                vn[9 * o + 3 * k + a] = nrm[q][a];
            }
            hs = k == 0 ? (double)p[q][2] : hs + (double)p[q][2];
        }
        const double hmean = hs / 3.0;
        const double span = hi - lo > 1e-12 ? hi - lo : 1e-12;
        const double tt = (hmean - lo) / span;
        int64_t idx = (int64_t)(tt * (double)texW);
        idx = idx < 0 ? 0 : (idx > (int64_t)texW - 1 ? (int64_t)texW - 1 : idx);
        const float u = ((float)idx + 0.5f) / (float)texW;
        for (int k = 0; k < 3; ++k) {
            uv[6 * o + 2 * k] = u;
            uv[6 * o + 2 * k + 1] = 0.5f;
        }
    }
}

// Min and max of a window of the grid: one block per row, atomics on the
// order-preserving integer images of the floats.
// (negative floats map below every non-negative one, in order; the map is
// its own inverse on the bit patterns)
__device__ __forceinline__ int32_t ordered(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : (int32_t)(0x80000000u - (uint32_t)b - 1u);
}
__global__ void k_range(const float* __restrict__ H, uint32_t n1, uint32_t x0, uint32_t y0,
                        uint32_t w, uint32_t h, int32_t* __restrict__ out) {
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (uint32_t r = blockIdx.x; r < h; r += gridDim.x)
        for (uint32_t q = threadIdx.x; q < w; q += blockDim.x) {
            const int32_t v = ordered(H[(uint64_t)(y0 + r) * n1 + x0 + q]);
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    atomicMin(&out[0], lo);
    atomicMax(&out[1], hi);
}
float from_ordered(int32_t o) {
    const int32_t b = o >= 0 ? o : (int32_t)(0x80000000u - (uint32_t)o - 1u);
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

__global__ void k_sample(const float* __restrict__ H, uint32_t n1, double step,
                         const float* __restrict__ xy, uint32_t count, float* __restrict__ z) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    // the nearest grid vertex
    double fx = (double)xy[2 * k] / step, fy = (double)xy[2 * k + 1] / step;
    int64_t i = (int64_t)floor(fx + 0.5), j = (int64_t)floor(fy + 0.5);
    i = i < 0 ? 0 : (i > (int64_t)n1 - 1 ? (int64_t)n1 - 1 : i);
    j = j < 0 ? 0 : (j > (int64_t)n1 - 1 ? (int64_t)n1 - 1 : j);
    z[k] = H[(uint64_t)j * n1 + (uint64_t)i];
}

unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int vox_terrain_heights(float* out, uint32_t n1, double extent, uint64_t seed_fbm,
                        uint64_t seed_ridged, double ridged_amp) {
    t_err.clear();
    if (!out || n1 < 2) return fail("vox_terrain_heights: null output or fewer than 2 vertices"), 0;
    if ((uint64_t)n1 * n1 >= (1ull << 40)) return fail("vox_terrain_heights: grid too large"), 0;
    Perms P;
    permutation(seed_fbm, P.fbm);
    permutation(seed_ridged, P.ridged);
    const double step = extent / (double)(n1 - 1);
    const uint64_t n = (uint64_t)n1 * n1;
    k_heights<<<blocks_for(n), 256>>>(out, n1, extent, step, ridged_amp, P);
    if (!check(hipGetLastError(), "vox_terrain_heights") ||
        !check(hipDeviceSynchronize(), "vox_terrain_heights"))
        return 0;
    return 1;
}

int vox_terrain_range(const float* heights, uint32_t n1, uint32_t x0, uint32_t y0, uint32_t w,
                      uint32_t h, float out_minmax[2]) {
    t_err.clear();
    if (!heights || !out_minmax || w == 0 || h == 0 || (uint64_t)x0 + w > n1 ||
        (uint64_t)y0 + h > n1)
        return fail("vox_terrain_range: window outside the grid"), 0;
    // the result words: one pair per calling thread, allocated on its first
    // call and kept (no hipFree, which would wait for the whole device)
    thread_local int32_t* d = nullptr;
    if (!d && !check(hipMalloc(&d, 8), "vox_terrain_range")) return 0;
    const int32_t init[2] = {INT32_MAX, INT32_MIN};
    int32_t r[2] = {0, 0};
    // on the null stream, which the library's non-blocking streams do not
    // wait for and which waits for none of them
    bool ok = check(hipMemcpyAsync(d, init, 8, hipMemcpyHostToDevice, nullptr), "vox_terrain_range");
    if (ok) {
        k_range<<<h < 4096u ? h : 4096u, 256>>>(heights, n1, x0, y0, w, h, d);
        ok = check(hipGetLastError(), "vox_terrain_range") &&
             check(hipMemcpyAsync(r, d, 8, hipMemcpyDeviceToHost, nullptr), "vox_terrain_range") &&
             check(hipStreamSynchronize(nullptr), "vox_terrain_range");
    }
    if (!ok) return 0;
    out_minmax[0] = from_ordered(r[0]);
    out_minmax[1] = from_ordered(r[1]);
    return 1;
}

int vox_terrain_tile(const float* heights, uint32_t n1, double extent, uint32_t x0, uint32_t y0,
                     uint32_t cw, uint32_t ch, float hmin, float hmax, uint32_t tex_width,
                     float* verts, float* vnormals, float* uvs) {
    t_err.clear();
    if (!heights || !verts || !vnormals || !uvs || tex_width == 0)
        return fail("vox_terrain_tile: null pointer or zero texture width"), 0;
    if (cw == 0 || ch == 0 || (uint64_t)x0 + cw >= n1 || (uint64_t)y0 + ch >= n1)
        return fail("vox_terrain_tile: cell window outside the grid"), 0;
    const double step = extent / (double)(n1 - 1);
    const float dxf = (float)step;
    const uint64_t n = (uint64_t)cw * ch;
    k_tile<<<blocks_for(n), 256>>>(heights, n1, step, dxf, x0, y0, cw, ch, (double)hmin,
                                   (double)hmax, tex_width, verts, vnormals, uvs);
    // waits for this launch only (the null stream), not for the device
    if (!check(hipGetLastError(), "vox_terrain_tile") ||
        !check(hipStreamSynchronize(nullptr), "vox_terrain_tile"))
        return 0;
    return 1;
}

int vox_terrain_sample(const float* heights, uint32_t n1, double extent, const float* xy,
                       uint32_t count, float* z) {
    t_err.clear();
    if (!heights || (count && (!xy || !z))) return fail("vox_terrain_sample: null pointer"), 0;
    if (count == 0) return 1;
    float *dxy = nullptr, *dz = nullptr;
    bool ok = check(hipMalloc(&dxy, 8ull * count), "vox_terrain_sample") &&
              check(hipMalloc(&dz, 4ull * count), "vox_terrain_sample") &&
              check(hipMemcpy(dxy, xy, 8ull * count, hipMemcpyHostToDevice), "vox_terrain_sample");
    if (ok) {
        k_sample<<<blocks_for(count), 256>>>(heights, n1, extent / (double)(n1 - 1), dxy, count, dz);
        ok = check(hipGetLastError(), "vox_terrain_sample") &&
             check(hipMemcpy(z, dz, 4ull * count, hipMemcpyDeviceToHost), "vox_terrain_sample");
    }
    if (dxy) (void)hipFree(dxy);
    if (dz) (void)hipFree(dz);
    return ok ? 1 : 0;
}

const char* vox_terrain_get_error(void) { return t_err.c_str(); }

}  // extern "C"
