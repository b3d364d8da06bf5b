// vox_serve.hip -- sparse-octree ray cast producing brick requests, and the
// request deduplication (C ABI in include/voxmi_serve.h).
//
// The ray cast restates VoxViz/Shaders/GigaVoxels.frag main (544-805) per
// pixel: descend from the root while the node is constant or has a resident
// brick, its projected voxel is larger than a pixel and it has children;
// at the stopping node composite its constant colour or the most recently
// descended resident brick up to the node's exit, record the node in the
// pixel's usage list, restart from the root.  fp32, no contraction, the
// shader's operation order; GLSL built-ins are defined as follows (the same
// definitions hold in the oracle):
//   normalize(v) = v * (1 / sqrt(dot(v, v))), dot = (x*x + y*y) + z*z
//   pow(x, 2^n)  = x squared n times (all exponents in the shader are 2^n)
//   texture()    = trilinear, clamp to edge, weights rounded to 8 fractional
//                  bits, unorm8 texels as c / 255
//   uvec3(v)     = truncation, negative components 0
// Rays start on the box of the octree (the rasterized front faces of the
// reference), at the camera when it is inside the box.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/voxmi_serve.h"
#include "vox_internal.h"

using namespace voxmi;

namespace {

constexpr int kMaxUsage = 12;  // 3 lists x 4 components per pixel
constexpr uint32_t kConstBit = 0x40000000u;
constexpr uint32_t kResidentBit = 0x80000000u;

struct RayUniforms {
    float camPos[3], camUp[3], camLeft[3];
    float volMin[3], volMax[3], volScale[3];
    float look[3], right[3], upv[3];  // world ray basis: look, right*tanX, up*tanY
    float rayStep, rootVoxelHalf, pixelSize, lodScalar;
    float brickStep[3], brickDim[3], poolDim[3];
    int32_t depthMinusOne, maxPush, W, H, maxTraversals;
    int32_t rx0, ry0, rx1, ry1;  // pixels cast: [rx0, rx1) x [ry0, ry1)
    uint32_t ndim[3], pdim[3];
    // wide brick pool: word 1 of a resident node is its slot, and each brick
    // is sampled in its own (B+2)^3 texel frame (poolDim / pdim are then the
    // brick's texels per axis)
    uint32_t wide, wideSlots;
};

struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v3 add(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 mul(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ v3 divv(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ v3 muls(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ v3 normalize(v3 v) { return muls(v, 1.0f / sqrtf(dot(v, v))); }
__device__ __forceinline__ float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ v3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ float inside_count(v3 p, v3 mn, v3 mx) {
    const v3 t1 = mk(signf(p.x - mn.x), signf(p.y - mn.y), signf(p.z - mn.z));
    const v3 t2 = mk(signf(mx.x - p.x), signf(mx.y - p.y), signf(mx.z - p.z));
    return dot(t1, t2);
}

// x^(2^n) by squaring.
__device__ __forceinline__ float pow2n(float x, int n) {
    for (int i = 0; i < n; ++i) x = x * x;
    return x;
}

// ComputeMinRayBoxIntersection (GigaVoxels.frag 155-180): the smallest
// positive-direction exit time over the six faces, 4 if none.
__device__ __forceinline__ float exit_time(v3 p, v3 d, v3 mn, v3 mx) {
    float tMin = 4.0f;
    // faces with outward normals -x, -y, -z through mn; +x, +y, +z through mx
    const float dd[6] = {-d.x, -d.y, -d.z, d.x, d.y, d.z};
    const float nv[6] = {-(p.x - mn.x), -(p.y - mn.y), -(p.z - mn.z),
                         p.x - mx.x, p.y - mx.y, p.z - mx.z};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (dd[i] > 0.0f) {
            const float t = -nv[i] / dd[i];
            if (t < tMin && fabsf(t) > 0.0f) tMin = t;
        }
    }
    return tMin;
}

__device__ __forceinline__ void unpack_rgba(uint32_t c, float o[4]) {
    o[0] = (float)(c & 255u) / 255.0f;
    o[1] = (float)((c >> 8) & 255u) / 255.0f;
    o[2] = (float)((c >> 16) & 255u) / 255.0f;
    o[3] = (float)(c >> 24) / 255.0f;
}

// texture(BrickSampler, s): trilinear, clamp to edge.  unorm8 channel c
// reads u8[c] = c / 255 (a table of the correctly rounded quotients); the
// pool has at most 2^30 texels, so texel indices fit 32 bits.  The eight
// texels are loaded before any is used (one memory round trip per sample),
// then weighted in corner order 0..7.
// A node texture or brick pool in device memory.  Typed so, loads through
// it are global loads also where the pointer itself was read from memory (an
// octree table), which the compiler would otherwise load with flat ones.
typedef const __attribute__((address_space(1))) uint32_t* gu32;

__device__ __forceinline__ void sample_brick(gu32 __restrict__ col,
                                             const RayUniforms& U,
                                             const float* __restrict__ u8, v3 s, float r[4]) {
    float c[3] = {s.x, s.y, s.z};
    int i0[3];
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float u = c[a];
        if (!(u == u)) u = 0.0f;
        const float b = u * U.poolDim[a] - 0.5f;
        const float f = floorf(b);
        w[a] = rintf((b - f) * 256.0f) / 256.0f;
        i0[a] = (int)fmaxf(fminf(f, 1.0e9f), -1.0e9f);
    }
    uint32_t t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        int ix = i0[0] + dx, iy = i0[1] + dy, iz = i0[2] + dz;
        ix = ix < 0 ? 0 : (ix >= (int)U.pdim[0] ? (int)U.pdim[0] - 1 : ix);
        iy = iy < 0 ? 0 : (iy >= (int)U.pdim[1] ? (int)U.pdim[1] - 1 : iy);
        iz = iz < 0 ? 0 : (iz >= (int)U.pdim[2] ? (int)U.pdim[2] - 1 : iz);
        t[k] = col[((uint32_t)iz * U.pdim[1] + (uint32_t)iy) * U.pdim[0] + (uint32_t)ix];
    }
    r[0] = r[1] = r[2] = r[3] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const float wx = dx ? w[0] : 1.0f - w[0];
        const float wy = dy ? w[1] : 1.0f - w[1];
        const float wz = dz ? w[2] : 1.0f - w[2];
        const float wk = (wx * wy) * wz;
        const uint32_t v = t[k];
        r[0] = r[0] + wk * u8[v & 255u];
        r[1] = r[1] + wk * u8[(v >> 8) & 255u];
        r[2] = r[2] + wk * u8[(v >> 16) & 255u];
        r[3] = r[3] + wk * u8[v >> 24];
    }
}

__device__ __forceinline__ void blend(float dst[4], const float src[4]) {
    const float k = 1.0f - dst[3];
    dst[0] = k * src[0] + dst[0];
    dst[1] = k * src[1] + dst[1];
    dst[2] = k * src[2] + dst[2];
    dst[3] = k * src[3] + dst[3];
}

__device__ __forceinline__ uint2 fetch_node(gu32 __restrict__ tex, const RayUniforms& U,
                                            uint32_t x, uint32_t y, uint32_t z) {
    // texelFetch outside the texture reads 0 (never reached with valid pointers)
    if (x >= U.ndim[0] || y >= U.ndim[1] || z >= U.ndim[2]) return make_uint2(0u, 0u);
    const uint64_t t = ((uint64_t)z * U.ndim[1] + y) * U.ndim[0] + x;
    return make_uint2(tex[2 * t], tex[2 * t + 1]);
}

// One pixel's ray through one octree (GigaVoxels.frag main): its FragColor
// in dst and its node usage list in usage (both start at zero here).
// Returns whether the ray met the octree's box.
__device__ __forceinline__ bool cast_pixel(const RayUniforms& U, gu32 __restrict__ tex,
                                           gu32 __restrict__ pool,
                                           const float* __restrict__ jitter,
                                           const float* __restrict__ u8, int i, int j,
                                           float (&dst)[4], uint32_t (&usage)[kMaxUsage]) {
#pragma unroll
    for (int k = 0; k < kMaxUsage; ++k) usage[k] = 0u;
    dst[0] = dst[1] = dst[2] = dst[3] = 0.0f;

    // the pixel's ray, world -> unit volume
    const float sx = (2.0f * ((float)i + 0.5f)) / (float)U.W - 1.0f;
    const float sy = (2.0f * ((float)j + 0.5f)) / (float)U.H - 1.0f;
    const v3 dw = add(add(ld3(U.look), muls(ld3(U.right), sx)), muls(ld3(U.upv), sy));
    const v3 dn = divv(dw, ld3(U.volScale));
    const v3 cam = ld3(U.camPos), vmin = ld3(U.volMin), vmax = ld3(U.volMax);
    float tEnter = -INFINITY, tExit = INFINITY;
    bool hit = true;
    const float dc[3] = {dn.x, dn.y, dn.z}, cc[3] = {cam.x, cam.y, cam.z};
    for (int a = 0; a < 3; ++a) {
        if (dc[a] == 0.0f) {
            if (cc[a] < U.volMin[a] || cc[a] > U.volMax[a]) hit = false;
            continue;
        }
        const float t0 = (U.volMin[a] - cc[a]) / dc[a];
        const float t1 = (U.volMax[a] - cc[a]) / dc[a];
        tEnter = fmaxf(tEnter, fminf(t0, t1));
        tExit = fminf(tExit, fmaxf(t0, t1));
    }
    const float tStart = fmaxf(tEnter, 0.0f);
    if (!(tExit >= tStart)) hit = false;

    if (hit) {
        v3 rayPos = add(cam, muls(dn, tStart));
        const v3 rayDir = normalize(dn);
        const float jit = jitter[(j & 31) * 32 + (i & 31)];
        rayPos = add(rayPos, muls(muls(rayDir, U.rayStep), jit));
        float inside = inside_count(rayPos, vmin, vmax);
        if (inside >= 3.0f) {
            int curList = 0;  // list * 4 + component
            const int startPush = (i % 2) * kMaxUsage + (j % 2) * kMaxUsage * 2;
            const int stopPush = startPush + kMaxUsage;
            const int maxPerBlock = kMaxUsage * 4;
            const v3 brickRayStep = mul(rayDir, ld3(U.brickStep));
            const v3 camUp = ld3(U.camUp), camLeft = ld3(U.camLeft);
            const uint2 root = fetch_node(tex, U, 0, 0, 0);
            uint32_t px = 0, py = 0, pz = 0;
            uint2 node = root;
            v3 extMin = vmin, extMax = vmax;
            float nodeSize = 1.0f, vh = U.rootVoxelHalf;
            v3 rayBrickPos = rayPos, curBrickRayPos = rayPos;
            uint2 curBrickPtr = root;
            const bool rootConst = (root.x & kConstBit) != 0u;  // RootNodeIsConstant
            bool loaded = false, curConst = rootConst;
            int curDepth = 0, curBrickDepth = 0;
            uint32_t cx = 0, cy = 0, cz = 0;
            int nlt = 0, pushed = 0, traversals = 0;
            while (inside >= 3.0f && traversals < U.maxTraversals) {
                curConst = (node.x & kConstBit) != 0u;
                loaded = (node.x & kResidentBit) != 0u;
                if (loaded) {
                    curBrickPtr = node;
                    curBrickRayPos = rayBrickPos;
                    curBrickDepth = curDepth;
                }
                // ComputeVoxelSize (GigaVoxels.frag 483-490) * LodScalar
                const v3 ll = normalize(sub(add(sub(rayPos, muls(camUp, vh)), muls(camLeft, vh)), cam));
                const v3 ur = normalize(sub(sub(add(rayPos, muls(camUp, vh)), muls(camLeft, vh)), cam));
                const float proj = (1.0f - dot(ll, ur)) * U.lodScalar;
                bool stop = (!curConst && !loaded) || U.pixelSize > proj;
                if (!stop) {
                    cx = (node.x >> 20) & 1023u;
                    cy = (node.x >> 10) & 1023u;
                    cz = node.x & 1023u;
                    // the depth bound never binds for a tree built here (the
                    // last depth has no child pointers); it keeps a damaged
                    // texture from descending forever
                    stop = (cx | cy | cz) == 0u || curDepth >= U.depthMinusOne;
                }
                if (stop) {
                    const uint32_t id = ((px + 1u) << 20) | ((py + 1u) << 10) | (pz + 1u);
                    if (curConst) {
                        const int n = U.depthMinusOne - curDepth;
                        const float ratio = ldexpf(1.0f, n);
                        const float step = U.rayStep * ratio;
                        float cv[4];
                        unpack_rgba(node.y, cv);
                        const float tExitNode = exit_time(rayPos, rayDir, extMin, extMax);
                        if (cv[3] < 0.005f) {
                            const int ns = (int)(tExitNode / step) + 1;  // AdvanceRay
                            rayPos = add(rayPos, muls(rayDir, (float)ns * step));
                        } else {
                            int ns = (int)(tExitNode / step) + 1;
                            const v3 ws = muls(rayDir, step);
                            float src[4];
                            src[3] = 1.0f - pow2n(1.0f - cv[3], n);
                            src[0] = cv[0] * ratio;
                            src[1] = cv[1] * ratio;
                            src[2] = cv[2] * ratio;
                            for (int s = 0; s < ns; ++s) {
                                rayPos = add(rayPos, ws);
                                if (dst[3] > 0.995f) {
                                    const int rest = ns - (s + 1);
                                    rayPos = add(rayPos, muls(rayDir, (float)rest * step));
                                    break;
                                }
                                blend(dst, src);
                            }
                        }
                    } else {
                        // the brick's first interior texel (slot origin + 1), in the
                        // pool's normalised coordinates or, in a wide pool, the brick's
                        const v3 bp = U.wide
                            ? mk(1.0f / U.poolDim[0], 1.0f / U.poolDim[1], 1.0f / U.poolDim[2])
                            : mk((float)((curBrickPtr.y >> 20) & 1023u) / U.poolDim[0],
                                 (float)((curBrickPtr.y >> 10) & 1023u) / U.poolDim[1],
                                 (float)(curBrickPtr.y & 1023u) / U.poolDim[2]);
                        // (a path with no resident brick keeps the root's word 1,
                        // which need not be a slot: slot 0 is read then, where the
                        // reference reads whatever texel those bits name)
                        const uint32_t slot = curBrickPtr.y < U.wideSlots ? curBrickPtr.y : 0u;
                        gu32 bpool =
                            U.wide ? pool + (uint64_t)slot * U.pdim[0] * U.pdim[1] * U.pdim[2] : pool;
                        v3 cbr = add(bp, divv(mul(curBrickRayPos, ld3(U.brickDim)),
                                              ld3(U.poolDim)));
                        const float ratio = ldexpf(1.0f, U.depthMinusOne - curBrickDepth);
                        const v3 ws = muls(muls(rayDir, U.rayStep), ratio);
                        const float tExitNode = exit_time(rayPos, rayDir, extMin, extMax);
                        const float step = U.rayStep * ratio;
                        int ns = (int)(tExitNode / step) + 1;
                        for (int s = 0; s < ns; ++s) {
                            rayPos = add(rayPos, ws);
                            if (dst[3] > 0.995f) {
                                const int rest = ns - (s + 1);
                                rayPos = add(rayPos, muls(rayDir, (float)rest * step));
                                break;
                            }
                            float src[4];
                            sample_brick(bpool, U, u8, cbr, src);
                            src[3] = 1.0f - (1.0f - src[3]);  // pow(1 - a, 1.0)
                            blend(dst, src);
                            cbr = add(cbr, brickRayStep);
                        }
                    }
                    if (nlt >= startPush && nlt < stopPush && pushed < U.maxPush) {
                        ++pushed;
#pragma unroll
                        for (int k = 0; k < kMaxUsage; ++k)
                            if (k == curList) usage[k] = id;
                        curList = curList + 1 == kMaxUsage ? 0 : curList + 1;
                    }
                    if (dst[3] > 0.995f && (curConst || loaded)) break;
                    // octTreeNodePtr is not reset (GigaVoxels.frag 718-742): a
                    // traversal that stops at the root records the previous
                    // stopping node's texel.
                    node = root;
                    extMin = vmin;
                    extMax = vmax;
                    nodeSize = 1.0f;
                    vh = U.rootVoxelHalf;
                    rayBrickPos = rayPos;
                    curBrickRayPos = rayPos;
                    curBrickPtr = node;
                    loaded = false;
                    curConst = rootConst;
                    curDepth = 0;
                    curBrickDepth = 0;
                    cx = cy = cz = 0;
                    inside = inside_count(rayPos, vmin, vmax);
                    ++traversals;
                    ++nlt;
                    if (nlt >= maxPerBlock) nlt = 0;
                } else {
                    const v3 bo = muls(rayBrickPos, 2.0f);
                    const uint32_t ox = bo.x >= 1.0f ? 1u : 0u;
                    const uint32_t oy = bo.y >= 1.0f ? 1u : 0u;
                    const uint32_t oz = bo.z >= 1.0f ? 1u : 0u;
                    rayBrickPos = sub(bo, mk((float)ox, (float)oy, (float)oz));
                    px = cx + ox;
                    py = cy + oy;
                    pz = cz + oz;
                    ++curDepth;
                    nodeSize *= 0.5f;
                    extMin = add(extMin, muls(mk((float)ox, (float)oy, (float)oz), nodeSize));
                    extMax = sub(extMax, muls(mk(1.0f - (float)ox, 1.0f - (float)oy,
                                                 1.0f - (float)oz), nodeSize));
                    vh *= 0.5f;
                    node = fetch_node(tex, U, px, py, pz);
                }
            }
        }
    }
    return hit;
}

__global__ void __launch_bounds__(256)
k_raycast(RayUniforms U, const uint32_t* __restrict__ tex, const uint32_t* __restrict__ pool,
          const float* __restrict__ jitter, uint32_t* __restrict__ rgbaOut,
          float4* __restrict__ rgbaF, uint32_t* __restrict__ usageOut) {
    __shared__ float u8[256];
    u8[threadIdx.x] = (float)threadIdx.x / 255.0f;  // blockDim.x == 256
    __syncthreads();
    const int i = U.rx0 + (int)blockIdx.x * 16 + (int)(threadIdx.x & 15);
    const int j = U.ry0 + (int)blockIdx.y * 16 + (int)(threadIdx.x >> 4);
    if (i >= U.rx1 || j >= U.ry1) return;
    uint32_t usage[kMaxUsage];
    float dst[4];
    cast_pixel(U, (gu32)tex, (gu32)pool, jitter, u8, i, j, dst, usage);
    const uint64_t pix = (uint64_t)j * U.W + i;
    if (rgbaOut) {
        uint32_t o = 0;
        for (int c = 0; c < 4; ++c) {
            float v = dst[c] * 255.0f;
            v = v > 255.0f ? 255.0f : (v > 0.0f ? v : 0.0f);
            o |= (uint32_t)v << (8 * c);
        }
        rgbaOut[pix] = o;
    }
    if (rgbaF) rgbaF[pix] = make_float4(dst[0], dst[1], dst[2], dst[3]);  // FragColor
    uint4* u4 = reinterpret_cast<uint4*>(usageOut + pix * kMaxUsage);
    u4[0] = make_uint4(usage[0], usage[1], usage[2], usage[3]);
    u4[1] = make_uint4(usage[4], usage[5], usage[6], usage[7]);
    u4[2] = make_uint4(usage[8], usage[9], usage[10], usage[11]);
}

// One octree of a front-to-back frame over several octrees: its uniforms
// and buffers, and the request list its rays append to.
struct TreeCast {
    RayUniforms U;
    const uint32_t* tex;
    const uint32_t* pool;
    uint32_t* stamp;              // node texel -> tag of the frame that listed it
    uint32_t* list;
    unsigned long long* times;
    uint32_t* counts;             // appended (capped), distinct
    uint32_t tag, cap;
};

// Append a wave's node ids of one usage slot to the octree's request list:
// lanes holding the same id elect one leader (ballot loop), leaders claim
// the node texel's frame stamp, the first claimant appends the id.
__device__ __forceinline__ void append_requests(const TreeCast& tc, uint32_t id) {
    uint64_t pending = __ballot(id != 0u);
    bool leader = false;
    while (pending) {
        const int l = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lid = __shfl(id, l);
        const uint64_t same = __ballot(id == lid);
        if ((int)(threadIdx.x & 63) == l) leader = true;
        pending &= ~same;
    }
    if (!leader) return;
    const uint32_t x = ((id >> 20) & 1023u) - 1u;
    const uint32_t y = ((id >> 10) & 1023u) - 1u;
    const uint32_t z = (id & 1023u) - 1u;
    const uint32_t* nd = tc.U.ndim;
    if (x >= nd[0] || y >= nd[1] || z >= nd[2]) return;
    const uint64_t t = ((uint64_t)z * nd[1] + y) * nd[0] + x;
    if (tc.stamp[t] != tc.tag && atomicExch(&tc.stamp[t], tc.tag) != tc.tag) {
        const uint32_t k = atomicAdd(&tc.counts[1], 1u);
        if (k < tc.cap) {
            tc.list[k] = id;
            tc.times[k] = (unsigned long long)wall_clock64();  // emitted
            atomicAdd(&tc.counts[0], 1u);
        }
    }
}

// Front to back over n octrees listed in an order every ray meets them in
// (boxes that tile space; the caller's order): each pixel casts its ray
// through every octree whose pixel rectangle holds it, exactly as one
// octree's frame does (cast_pixel), and accumulates the casts' FragColors
// as the renderer's far-to-near blend would (glBlendFunc(GL_SRC_ALPHA,
// GL_ONE_MINUS_SRC_ALPHA) per channel, alpha included): out = sum over
// octrees of T * src * src.a, T the product of (1 - src.a) of the octrees
// before.  A pixel stops once T <= tStop (tStop 0: never, not even at
// T = 0), a wave once all its pixels have; the octrees it skips list no
// requests for it.  With resume the pass continues a frame: acc from frame,
// T from trans (a pass over the octrees in front wrote both); trans, when
// given, receives each pixel's T for the next pass.
__global__ void __launch_bounds__(256)
k_raycast_multi(const TreeCast* __restrict__ trees, uint32_t n, const float* __restrict__ jitter,
                int32_t W, int32_t x0, int32_t y0, int32_t x1, int32_t y1, float tStop,
                float4* __restrict__ frame, float* __restrict__ trans, int resume) {
    __shared__ float u8[256];
    u8[threadIdx.x] = (float)threadIdx.x / 255.0f;  // blockDim.x == 256
    __syncthreads();
    // pixels of the union of the octrees' rectangles, [x0, x1) x [y0, y1)
    const int i = x0 + (int)blockIdx.x * 16 + (int)(threadIdx.x & 15);
    const int j = y0 + (int)blockIdx.y * 16 + (int)(threadIdx.x >> 4);
    const bool in = i < x1 && j < y1;
    const uint64_t pix = (uint64_t)j * W + i;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float T = 1.0f;
    if (resume && in) {   // the frame so far, from the octrees in front
        const float4 f = frame[pix];
        acc[0] = f.x;
        acc[1] = f.y;
        acc[2] = f.z;
        acc[3] = f.w;
        T = trans[pix];
    }
    for (uint32_t k = 0; k < n; ++k) {
        const bool want = in && (tStop <= 0.0f || T > tStop);
        if (!__any(want)) break;
        const TreeCast& tc = trees[k];
        const RayUniforms& U = tc.U;
        uint32_t usage[kMaxUsage];
        float dst[4];
        bool act = want && i >= U.rx0 && i < U.rx1 && j >= U.ry0 && j < U.ry1;
        if (act) {
            act = cast_pixel(U, (gu32)tc.tex, (gu32)tc.pool, jitter, u8, i, j, dst, usage);
        } else {
#pragma unroll
            for (int q = 0; q < kMaxUsage; ++q) usage[q] = 0u;
        }
        if (act) {
            const float a = dst[3];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = acc[c] + T * (dst[c] * a);
            T = T * (1.0f - a);
        }
#pragma unroll
        for (int q = 0; q < kMaxUsage; ++q) append_requests(tc, usage[q]);
    }
    if (in) {
        frame[pix] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        if (trans) trans[pix] = T;
    }
}


// Distinct node ids of the frame's pixel rectangle: lanes of a wave
// holding the same id elect one leader (ballot loop), leaders claim the node
// texel's frame stamp, the first claimant appends the id.  counts[0] =
// appended (capped), counts[1] = distinct ids.  Entry e of the rectangle is
// list slot e % 12 of its pixel e / 12 (row by row).
__global__ void __launch_bounds__(256)
k_requests(const uint32_t* __restrict__ usage, uint32_t W, int32_t rx0, int32_t ry0, uint32_t rw,
           uint64_t n, uint32_t tag, uint32_t ndx, uint32_t ndy, uint32_t ndz,
           uint32_t* __restrict__ stamp, uint32_t* __restrict__ list,
           unsigned long long* __restrict__ times, uint32_t cap,
           uint32_t* __restrict__ counts) {
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n;
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = base + threadIdx.x;
        uint32_t id = 0u;
        if (e < n) {
            // n < 2^32 (checked at launch): 32-bit index arithmetic
            const uint32_t e32 = (uint32_t)e;
            const uint32_t p = e32 / (uint32_t)kMaxUsage;
            const uint32_t k = e32 - p * (uint32_t)kMaxUsage;
            const uint32_t prow = p / rw;
            const uint64_t px = (uint64_t)rx0 + (p - prow * rw), py = (uint64_t)ry0 + prow;
            id = usage[(py * W + px) * kMaxUsage + k];
        }
        // wave-level deduplication
        uint64_t pending = __ballot(id != 0u);
        bool leader = false;
        while (pending) {
            const int l = __ffsll((unsigned long long)pending) - 1;
            const uint32_t lid = __shfl(id, l);
            const uint64_t same = __ballot(id == lid);
            if ((int)(threadIdx.x & 63) == l) leader = true;
            pending &= ~same;
        }
        if (leader) {
            const uint32_t x = ((id >> 20) & 1023u) - 1u;
            const uint32_t y = ((id >> 10) & 1023u) - 1u;
            const uint32_t z = (id & 1023u) - 1u;
            const uint64_t t = ((uint64_t)z * ndy + y) * ndx + x;
            // a plain read first: most leaders meet a node another wave has
            // claimed already, and skip the atomic on its (hot) stamp word
            if (x < ndx && y < ndy && z < ndz && stamp[t] != tag &&
                atomicExch(&stamp[t], tag) != tag) {
                const uint32_t k = atomicAdd(&counts[1], 1u);
                if (k < cap) {
                    list[k] = id;
                    times[k] = (unsigned long long)wall_clock64();  // emitted
                    atomicAdd(&counts[0], 1u);
                }
            }
        }
    }
}

// One caster's pixel rectangle blended into the frame:
// glBlendFunc(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA) on an RGBA32F target,
// out = src * src.a + dst * (1 - src.a) for all four channels.
__global__ void k_blend(const float4* __restrict__ src, int32_t W, int32_t x0, int32_t y0,
                        int32_t x1, int32_t y1, float4* __restrict__ dst) {
    const int32_t i = x0 + (int32_t)(blockIdx.x * 16u + (threadIdx.x & 15u));
    const int32_t j = y0 + (int32_t)(blockIdx.y * 16u + (threadIdx.x >> 4));
    if (i >= x1 || j >= y1) return;
    const uint64_t p = (uint64_t)j * W + i;
    const float4 s = src[p], d = dst[p];
    const float a = s.w, k = 1.0f - s.w;
    dst[p] = make_float4(s.x * a + d.x * k, s.y * a + d.y * k, s.z * a + d.z * k,
                         s.w * a + d.w * k);
}

// Jitter table: 32 x 32 values in [0, 1), the k-th from the k+1-th output of
// splitmix64 seeded with `seed` (the state advances by the golden-ratio
// increment, so output k+1 is the finaliser of seed + (k+1) * increment).
__global__ void k_jitter(uint64_t seed, float* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 1024) return;
    uint64_t z = seed + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[k] = (float)(z >> 40) * (1.0f / 16777216.0f);
}

}  // namespace

struct vox_raycaster {
    vox_octree* tree = nullptr;
    bool valid = false;
    int32_t maxW = 0, maxH = 0, W = 0, H = 0;
    uint32_t maxReq = 0, nreq = 0, nreqTotal = 0, tag = 0;
    uint32_t ndim[3] = {0, 0, 0};
    hipStream_t st = nullptr;  // the ray caster's own stream
    DevBuf rgba, rgbaF, usage, req, reqTime, counts, stamp, jitter;
    DevBuf multiTable;          // octree table of front-to-back frames queued here
    // its pinned host staging, two buffers in turn, each reused only after
    // the copy that read it last has run
    void* hostTable[2] = {nullptr, nullptr};
    size_t hostCap[2] = {0, 0};
    hipEvent_t tableCopied[2] = {nullptr, nullptr};
    int tableNext = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t done = nullptr;  // the last frame's request list is complete
    hipEvent_t read = nullptr;  // a composite has read the last frame
    int32_t rect[4] = {0, 0, 0, 0};  // pixels the last frame wrote
    bool timingPending = false, countsPending = false, haveJitter = false, readPending = false;
    uint64_t jitterSeed = 0;
    double msRay = 0, msReq = 0;
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

namespace voxmi {

bool raycaster_requests(vox_raycaster* rc, RequestRefs& r) {
    if (!rc || !rc->valid) return false;
    r.ids = rc->req.as<uint32_t>();
    r.times = rc->reqTime.as<uint64_t>();
    r.count = rc->counts.as<uint32_t>();
    r.maxIds = rc->maxReq;
    r.done = rc->done;
    r.tree = rc->tree;
    return true;
}

}  // namespace voxmi

namespace {

// Uniforms of one frame (GigaVoxelsRenderer.cpp 1650-1816), in float like
// the renderer's QVector3D arithmetic.
bool ray_uniforms(const vox_camera& cam, const vox_raycast_params& prm, int levels,
                  const uint32_t B[3], const uint32_t poolDims[3], uint32_t poolWide,
                  const uint32_t ndim[3], RayUniforms& U) {
    std::memset(&U, 0, sizeof(U));
    float T[3], S[3];
    for (int a = 0; a < 3; ++a) {
        T[a] = (prm.vol_min[a] + prm.vol_max[a]) * 0.5f;
        S[a] = prm.vol_max[a] - prm.vol_min[a];
        if (!(S[a] > 0.0f)) return false;
        U.volScale[a] = S[a];
        U.volMin[a] = (prm.vol_min[a] - T[a]) / S[a] + 0.5f;
        U.volMax[a] = (prm.vol_max[a] - T[a]) / S[a] + 0.5f;
        U.camPos[a] = (cam.position[a] - T[a]) / S[a] + 0.5f;
        U.camUp[a] = cam.up[a];
    }
    // left = up x look
    U.camLeft[0] = cam.up[1] * cam.look[2] - cam.up[2] * cam.look[1];
    U.camLeft[1] = cam.up[2] * cam.look[0] - cam.up[0] * cam.look[2];
    U.camLeft[2] = cam.up[0] * cam.look[1] - cam.up[1] * cam.look[0];
    const float tanY = std::tan(cam.fov_y_deg * 0.5f * 3.14159265358979f / 180.0f);
    const float tanX = tanY * (float)cam.width / (float)cam.height;
    for (int a = 0; a < 3; ++a) {
        U.look[a] = cam.look[a];
        U.right[a] = -U.camLeft[a] * tanX;
        U.upv[a] = cam.up[a] * tanY;
    }
    // RayStepSize: half the smallest voxel edge of the unit volume
    const uint32_t Nmax = std::max(B[0], std::max(B[1], B[2])) << (levels - 1);
    U.rayStep = prm.ray_step > 0.0f ? prm.ray_step : (1.0f / (float)Nmax) * 0.5f;
    if (!(U.rayStep >= 1.0e-5f)) return false;  // bounds the samples per node
    U.lodScalar = prm.lod_scalar > 0.0f ? prm.lod_scalar : 10.0f;
    // BrickStepVector = (1 / pool) * (rayStep * 2^(L-1) * S) / (S / B)
    const float pow2 = std::ldexp(1.0f, levels - 1);
    for (int a = 0; a < 3; ++a) {
        const float rayStepRoot = (S[a] * U.rayStep) * pow2;
        const float rootVoxel = S[a] / (float)B[a];
        const float ratio = rayStepRoot / rootVoxel;
        // a wide pool is sampled one brick at a time: its texels are the frame
        const uint32_t pd = poolWide ? B[a] + 2u : poolDims[a];
        U.brickStep[a] = (1.0f / (float)pd) * ratio;
        U.brickDim[a] = (float)B[a];
        U.poolDim[a] = (float)pd;
        U.pdim[a] = pd;
        U.ndim[a] = ndim[a];
    }
    U.wide = poolWide;
    U.wideSlots = poolWide ? poolDims[2] / (B[2] + 2u) : 0u;
    const float bx = (float)B[0], by = (float)B[1], bz = (float)B[2];
    const float minB = bx < ((by < bz) ? by : bz) ? bx : bz;  // renderer's ternary
    U.rootVoxelHalf = (1.0f / minB) * 0.5f;
    // PixelSize: 1 - cos of the angle between the rays through the centre
    // pixel corner and its diagonal neighbour (Camera::getPixelSize).
    {
        const double W = cam.width, H = cam.height;
        double r0[3], r1[3];
        for (int a = 0; a < 3; ++a) {
            const double sx0 = 2.0 * (W * 0.5) / W - 1.0, sy0 = 2.0 * (H * 0.5) / H - 1.0;
            const double sx1 = 2.0 * (W * 0.5 + 1.0) / W - 1.0, sy1 = 2.0 * (H * 0.5 + 1.0) / H - 1.0;
            r0[a] = cam.look[a] + (double)U.right[a] * sx0 + (double)U.upv[a] * sy0;
            r1[a] = cam.look[a] + (double)U.right[a] * sx1 + (double)U.upv[a] * sy1;
        }
        const double n0 = std::sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
        const double n1 = std::sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
        const double c = (r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2]) / (n0 * n1);
        U.pixelSize = (float)(1.0 - c);
    }
    U.depthMinusOne = levels - 1;
    U.maxPush = kMaxUsage * (int)((cam.frame % 2) + 1);
    U.W = cam.width;
    U.H = cam.height;
    const int32_t* r = prm.rect;
    if (r[0] == 0 && r[1] == 0 && r[2] == 0 && r[3] == 0) {
        U.rx0 = U.ry0 = 0;
        U.rx1 = cam.width;
        U.ry1 = cam.height;
    } else {
        if (r[0] < 0 || r[1] < 0 || r[2] > cam.width || r[3] > cam.height || r[2] <= r[0] ||
            r[3] <= r[1])
            return false;
        U.rx0 = r[0];
        U.ry0 = r[1];
        U.rx1 = r[2];
        U.ry1 = r[3];
    }
    U.maxTraversals = prm.max_traversals ? (int32_t)prm.max_traversals : 4096;
    return true;
}

}  // namespace

extern "C" {

vox_raycaster* vox_raycaster_create(vox_octree* tree, int32_t max_width, int32_t max_height,
                                    uint32_t max_requests) {
    vox_raycaster* rc = new (std::nothrow) vox_raycaster;
    if (!rc) return nullptr;
    rc->tree = tree;
    OctreeRefs o;
    if (!tree || !octree_refs(tree, o)) {
        rc->fail("raycaster: invalid octree");
        return rc;
    }
    for (int a = 0; a < 3; ++a) rc->ndim[a] = o.ndim[a];
    if (max_width <= 0 || max_height <= 0 || max_requests == 0) {
        rc->fail("raycaster: empty frame or request capacity");
        return rc;
    }
    if (!rc->check(hipStreamCreateWithFlags(&rc->st, hipStreamNonBlocking), "hipStreamCreate"))
        return rc;
    rc->maxW = max_width;
    rc->maxH = max_height;
    rc->maxReq = max_requests;
    const uint64_t npix = (uint64_t)max_width * max_height;
    const uint64_t ntex = (uint64_t)rc->ndim[0] * rc->ndim[1] * rc->ndim[2];
    if (!rc->rgba.ensure(4 * npix, rc->err, "rgba") ||
        !rc->rgbaF.ensure(16 * npix, rc->err, "rgba f32") ||
        !rc->usage.ensure(4 * npix * kMaxUsage, rc->err, "usage") ||
        !rc->req.ensure(4ull * max_requests, rc->err, "requests") ||
        !rc->reqTime.ensure(8ull * max_requests, rc->err, "request times") ||
        !rc->counts.ensure(8, rc->err, "counts") || !rc->stamp.ensure(4 * ntex, rc->err, "stamp") ||
        !rc->jitter.ensure(4 * 1024, rc->err, "jitter"))
        return rc;
    launch_fill_u32(rc->st, rc->stamp.as<uint32_t>(), ntex, 0u);
    launch_fill_u32(rc->st, rc->counts.as<uint32_t>(), 2, 0u);
    for (auto& e : rc->ev)
        if (!rc->check(hipEventCreate(&e), "hipEventCreate")) return rc;
    if (!rc->check(hipEventCreateWithFlags(&rc->read, hipEventDisableTiming), "hipEventCreate") ||
        !rc->check(hipEventCreateWithFlags(&rc->done, hipEventDisableTiming), "hipEventCreate") ||
        !rc->check(hipEventRecord(rc->done, rc->st), "event record") ||
        !rc->check(hipStreamSynchronize(rc->st), "synchronize"))
        return rc;
    rc->valid = true;
    return rc;
}

void vox_raycaster_destroy(vox_raycaster* rc) {
    if (!rc) return;
    if (rc->st) (void)hipStreamSynchronize(rc->st);
    for (auto e : rc->ev)
        if (e) (void)hipEventDestroy(e);
    if (rc->done) (void)hipEventDestroy(rc->done);
    if (rc->read) (void)hipEventDestroy(rc->read);
    rc->rgba.release();
    rc->rgbaF.release();
    rc->usage.release();
    rc->req.release();
    rc->reqTime.release();
    rc->counts.release();
    rc->stamp.release();
    rc->jitter.release();
    rc->multiTable.release();
    for (int b = 0; b < 2; ++b) {
        if (rc->hostTable[b]) (void)hipHostFree(rc->hostTable[b]);
        if (rc->tableCopied[b]) (void)hipEventDestroy(rc->tableCopied[b]);
    }
    if (rc->st) (void)hipStreamDestroy(rc->st);
    delete rc;
}

const char* vox_raycaster_get_error(vox_raycaster* rc) {
    if (!rc) return "null raycaster";
    rc->errCache = rc->err.str();
    return rc->errCache.c_str();
}

namespace {

// Timings of the frame whose events were recorded last (waits for it).
void settle_timing(vox_raycaster* rc) {
    if (!rc->timingPending) return;
    (void)hipEventSynchronize(rc->ev[2]);
    float a = 0, b = 0;
    (void)hipEventElapsedTime(&a, rc->ev[0], rc->ev[1]);
    (void)hipEventElapsedTime(&b, rc->ev[1], rc->ev[2]);
    rc->msRay = a;
    rc->msReq = b;
    rc->timingPending = false;
}

}  // namespace

int vox_raycast_frame(vox_raycaster* rc, const vox_camera* cam, const vox_raycast_params* prm) {
    if (!rc || !rc->valid || !cam || !prm) return 0;
    if (cam->width <= 0 || cam->height <= 0 || cam->width > rc->maxW || cam->height > rc->maxH)
        return rc->fail("raycast: viewport exceeds the raycaster's frame size"), 0;
    OctreeRefs o;
    if (!octree_refs(rc->tree, o)) return rc->fail("raycast: octree unavailable"), 0;
    const uint32_t* nd = o.ndim;
    uint32_t* tex = o.tex;
    hipStream_t st = rc->st;
    RayUniforms U;
    if (!ray_uniforms(*cam, *prm, o.levels, o.B, o.poolDims, o.poolWide, o.ndim, U))
        return rc->fail("raycast: empty volume box, ray step below 1e-5 or a pixel rectangle "
                        "outside the viewport"), 0;
    settle_timing(rc);  // the previous frame's events are reused below
    // the node texture and brick pool as the last writer (octree build or
    // brick server frame) left them
    if (!rc->check(hipStreamWaitEvent(st, o.ready, 0), "wait for the node texture")) return 0;
    // a composite still reading the last frame
    if (rc->readPending && !rc->check(hipStreamWaitEvent(st, rc->read, 0), "wait for a composite"))
        return 0;
    rc->readPending = false;
    if (!rc->haveJitter || rc->jitterSeed != prm->jitter_seed) {
        k_jitter<<<4, 256, 0, st>>>(prm->jitter_seed, rc->jitter.as<float>());
        rc->jitterSeed = prm->jitter_seed;
        rc->haveJitter = true;
    }
    rc->W = cam->width;
    rc->H = cam->height;
    const uint32_t rw = (uint32_t)(U.rx1 - U.rx0), rh = (uint32_t)(U.ry1 - U.ry0);
    const uint64_t npix = (uint64_t)rw * rh;
    if (npix * kMaxUsage >= (1ull << 32))
        return rc->fail("raycast: 2^32 usage entries or more"), 0;
    if (++rc->tag == 0) {  // stamp wrap: clear
        launch_fill_u32(st, rc->stamp.as<uint32_t>(), (uint64_t)nd[0] * nd[1] * nd[2], 0u);
        rc->tag = 1;
    }
    launch_fill_u32(st, rc->counts.as<uint32_t>(), 2, 0u);
    (void)hipEventRecord(rc->ev[0], st);
    dim3 grid((rw + 15) / 16, (rh + 15) / 16);
    k_raycast<<<grid, 256, 0, st>>>(U, tex, o.poolCol, rc->jitter.as<float>(), rc->rgba.as<uint32_t>(),
                                    rc->rgbaF.as<float4>(), rc->usage.as<uint32_t>());
    rc->rect[0] = U.rx0;
    rc->rect[1] = U.ry0;
    rc->rect[2] = U.rx1;
    rc->rect[3] = U.ry1;
    (void)hipEventRecord(rc->ev[1], st);
    const uint64_t nent = npix * kMaxUsage;
    const unsigned rgrid = (unsigned)std::min<uint64_t>((nent + 255) / 256, 8192);
    k_requests<<<rgrid, 256, 0, st>>>(rc->usage.as<uint32_t>(), (uint32_t)rc->W, U.rx0, U.ry0, rw,
                                      nent, rc->tag, nd[0], nd[1],
                                      nd[2], rc->stamp.as<uint32_t>(), rc->req.as<uint32_t>(),
                                      rc->reqTime.as<unsigned long long>(), rc->maxReq,
                                      rc->counts.as<uint32_t>());
    (void)hipEventRecord(rc->ev[2], st);
    if (!rc->check(hipGetLastError(), "raycast kernels") ||
        !rc->check(hipEventRecord(rc->done, st), "event record"))
        return 0;
    rc->timingPending = true;
    rc->countsPending = true;
    return 1;
}

int vox_raycast_frame_multi(vox_raycaster* const* rcs, uint32_t n, const vox_camera* cam,
                            const vox_raycast_params* prms, float* frame, float* transmittance,
                            float t_stop, int32_t resume) {
    if (!rcs || n == 0 || !cam || !prms || !frame || !(t_stop >= 0.0f)) return 0;
    if (resume && !transmittance)
        return rcs[0] ? rcs[0]->fail("raycast multi: resuming needs the transmittance buffer"),
               0 : 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (!rcs[k] || !rcs[k]->valid) return 0;
        for (uint32_t m = 0; m < k; ++m)
            if (rcs[m] == rcs[k]) return rcs[k]->fail("raycast multi: a ray caster appears twice"), 0;
    }
    vox_raycaster* r0 = rcs[0];
    if (cam->width <= 0 || cam->height <= 0) return r0->fail("raycast multi: empty viewport"), 0;
    hipStream_t st = r0->st;
    std::vector<TreeCast> table(n);
    for (uint32_t k = 0; k < n; ++k) {
        vox_raycaster* rc = rcs[k];
        if (cam->width > rc->maxW || cam->height > rc->maxH)
            return rc->fail("raycast multi: viewport exceeds the ray caster's frame size"), 0;
        if (prms[k].jitter_seed != prms[0].jitter_seed)
            return rc->fail("raycast multi: the octrees' jitter seeds differ"), 0;
        OctreeRefs o;
        if (!octree_refs(rc->tree, o)) return rc->fail("raycast multi: octree unavailable"), 0;
        TreeCast& tc = table[k];
        if (!ray_uniforms(*cam, prms[k], o.levels, o.B, o.poolDims, o.poolWide, o.ndim, tc.U))
            return rc->fail("raycast multi: empty volume box, ray step below 1e-5 or a pixel "
                            "rectangle outside the viewport"), 0;
        settle_timing(rc);
        // this octree's node texture and pool as their last writer left them,
        // and its own stream's earlier work
        if (!rc->check(hipStreamWaitEvent(st, o.ready, 0), "wait for the node texture") ||
            (rc != r0 && !rc->check(hipStreamWaitEvent(st, rc->done, 0), "wait for a frame")))
            return 0;
        if (rc->readPending && !rc->check(hipStreamWaitEvent(st, rc->read, 0), "wait for a composite"))
            return 0;
        rc->readPending = false;
        const uint64_t ntex = (uint64_t)rc->ndim[0] * rc->ndim[1] * rc->ndim[2];
        if (++rc->tag == 0) {  // stamp wrap: clear
            launch_fill_u32(st, rc->stamp.as<uint32_t>(), ntex, 0u);
            rc->tag = 1;
        }
        launch_fill_u32(st, rc->counts.as<uint32_t>(), 2, 0u);
        tc.tex = o.tex;
        tc.pool = o.poolCol;
        tc.stamp = rc->stamp.as<uint32_t>();
        tc.list = rc->req.as<uint32_t>();
        tc.times = rc->reqTime.as<unsigned long long>();
        tc.counts = rc->counts.as<uint32_t>();
        tc.tag = rc->tag;
        tc.cap = rc->maxReq;
        rc->W = cam->width;
        rc->H = cam->height;
        rc->rect[0] = tc.U.rx0;
        rc->rect[1] = tc.U.ry0;
        rc->rect[2] = tc.U.rx1;
        rc->rect[3] = tc.U.ry1;
    }
    if (!r0->haveJitter || r0->jitterSeed != prms[0].jitter_seed) {
        k_jitter<<<4, 256, 0, st>>>(prms[0].jitter_seed, r0->jitter.as<float>());
        r0->jitterSeed = prms[0].jitter_seed;
        r0->haveJitter = true;
    }
    // a new frame starts cleared (colour 0, transmittance 1); a pass then
    // covers the union of its octrees' rectangles only
    if (!resume) {
        launch_fill_u32(st, reinterpret_cast<uint32_t*>(frame), 4ull * cam->width * cam->height,
                        0u);
        if (transmittance)
            launch_fill_u32(st, reinterpret_cast<uint32_t*>(transmittance),
                            (uint64_t)cam->width * cam->height, 0x3f800000u);  // 1.0f
    }
    int32_t ux0 = cam->width, uy0 = cam->height, ux1 = 0, uy1 = 0;
    for (const TreeCast& tc : table) {
        ux0 = std::min(ux0, tc.U.rx0);
        uy0 = std::min(uy0, tc.U.ry0);
        ux1 = std::max(ux1, tc.U.rx1);
        uy1 = std::max(uy1, tc.U.ry1);
    }
    // the table is copied in stream order, after the previous frame's kernel
    // has read the device copy, from a pinned staging buffer whose last copy
    // has run
    const size_t tbytes = sizeof(TreeCast) * n;
    const int b = r0->tableNext;
    r0->tableNext ^= 1;
    if (!r0->tableCopied[b] &&
        !r0->check(hipEventCreateWithFlags(&r0->tableCopied[b], hipEventDisableTiming),
                   "hipEventCreate"))
        return 0;
    if (!r0->check(hipEventSynchronize(r0->tableCopied[b]), "octree table staging")) return 0;
    if (r0->hostCap[b] < tbytes) {
        if (r0->hostTable[b]) (void)hipHostFree(r0->hostTable[b]);
        r0->hostTable[b] = nullptr;
        r0->hostCap[b] = 0;
        if (!r0->check(hipHostMalloc(&r0->hostTable[b], tbytes, hipHostMallocDefault),
                       "pinned octree table"))
            return 0;
        r0->hostCap[b] = tbytes;
    }
    std::memcpy(r0->hostTable[b], table.data(), tbytes);
    if (!r0->multiTable.ensure(tbytes, r0->err, "octree table") ||
        !r0->check(hipMemcpyAsync(r0->multiTable.p, r0->hostTable[b], tbytes,
                                  hipMemcpyHostToDevice, st),
                   "octree table") ||
        !r0->check(hipEventRecord(r0->tableCopied[b], st), "event record"))
        return 0;
    (void)hipEventRecord(r0->ev[0], st);
    dim3 grid((unsigned)(ux1 - ux0 + 15) / 16, (unsigned)(uy1 - uy0 + 15) / 16);
    k_raycast_multi<<<grid, 256, 0, st>>>(r0->multiTable.as<TreeCast>(), n, r0->jitter.as<float>(),
                                          cam->width, ux0, uy0, ux1, uy1, t_stop,
                                          reinterpret_cast<float4*>(frame), transmittance,
                                          resume ? 1 : 0);
    (void)hipEventRecord(r0->ev[1], st);
    (void)hipEventRecord(r0->ev[2], st);
    if (!r0->check(hipGetLastError(), "raycast multi kernel")) return 0;
    r0->timingPending = true;
    for (uint32_t k = 0; k < n; ++k) {
        if (!rcs[k]->check(hipEventRecord(rcs[k]->done, st), "event record")) return 0;
        rcs[k]->countsPending = true;
    }
    return 1;
}

int vox_composite(vox_raycaster* const* rcs, uint32_t n, float* dst, int32_t width,
                  int32_t height, int32_t clear) {
    if (!dst || width <= 0 || height <= 0 || (n && !rcs)) return 0;
    for (uint32_t k = 0; k < n; ++k) {
        vox_raycaster* rc = rcs[k];
        if (!rc || !rc->valid) return 0;
        if (rc->W != width || rc->H != height)
            return rc->fail("composite: the frame size differs from the ray caster's last frame"), 0;
    }
    // the blends run in the given order on the first caster's stream (the
    // null stream when there is none), after every caster's frame
    hipStream_t st = n ? rcs[0]->st : nullptr;
    for (uint32_t k = 0; k < n; ++k)
        if (hipStreamWaitEvent(st, rcs[k]->done, 0) != hipSuccess) return 0;
    if (clear) launch_fill_u32(st, reinterpret_cast<uint32_t*>(dst), 4ull * width * height, 0u);
    for (uint32_t k = 0; k < n; ++k) {
        const int32_t* r = rcs[k]->rect;
        if (r[2] <= r[0] || r[3] <= r[1]) continue;
        dim3 grid((unsigned)(r[2] - r[0] + 15) / 16, (unsigned)(r[3] - r[1] + 15) / 16);
        k_blend<<<grid, 256, 0, st>>>(rcs[k]->rgbaF.as<float4>(), width, r[0], r[1], r[2], r[3],
                                      reinterpret_cast<float4*>(dst));
    }
    if (hipGetLastError() != hipSuccess) return 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (hipEventRecord(rcs[k]->read, st) != hipSuccess) return 0;
        rcs[k]->readPending = true;
    }
    if (!n && hipStreamSynchronize(st) != hipSuccess) return 0;
    return 1;
}

int vox_raycaster_get_frame(vox_raycaster* rc, vox_frame_view* out) {
    if (!rc || !rc->valid || !out) return 0;
    if (rc->countsPending) {
        uint32_t cnt[2] = {0, 0};
        // the frame may have been cast on another caster's stream (a
        // front-to-back pass); its done event marks the end either way
        if (!rc->check(hipStreamWaitEvent(rc->st, rc->done, 0), "wait for the frame") ||
            !rc->check(hipMemcpyAsync(cnt, rc->counts.p, 8, hipMemcpyDeviceToHost, rc->st),
                       "counts") ||
            !rc->check(hipStreamSynchronize(rc->st), "synchronize"))
            return 0;
        rc->nreq = cnt[0];
        rc->nreqTotal = cnt[1];
        rc->countsPending = false;
    }
    settle_timing(rc);
    out->width = rc->W;
    out->height = rc->H;
    out->rgba = rc->rgba.as<uint32_t>();
    out->rgba_f32 = rc->rgbaF.as<float>();
    out->usage = rc->usage.as<uint32_t>();
    out->requests = rc->req.as<uint32_t>();
    out->nrequests = rc->nreq;
    out->nrequests_total = rc->nreqTotal;
    out->ms_raycast = rc->msRay;
    out->ms_requests = rc->msReq;
    return 1;
}

}  // extern "C"
