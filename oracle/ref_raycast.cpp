// ref_raycast.cpp -- CPU restatement of the GigaVoxels ray cast that emits
// node-usage requests (VoxViz/Shaders/GigaVoxels.frag 1-805) and of the
// renderer's uniforms (VoxViz/GigaVoxels/GigaVoxelsRenderer.cpp 1650-1816).
//
// TEST INFRASTRUCTURE ONLY: the oracle for the HIP ray cast; only tests/ load
// it.  Written as the shader is, with a small vec3 type; GLSL built-ins take
// the definitions the HIP kernel documents (normalize = v * (1/sqrt(dot)),
// pow(x, 2^n) by squaring, trilinear texture with 8-bit weights and clamp to
// edge, uvec3 truncating with negatives at 0).  Rays start where they enter
// the volume box (the rasterized front faces), or at the camera inside it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <vector>

#include "ref_raycast.h"

namespace {

struct vec3 {
    float x, y, z;
    vec3() : x(0), y(0), z(0) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit vec3(float a) : x(a), y(a), z(a) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(vec3 a, vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
vec3 normalize(vec3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
float gsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
vec3 sign(vec3 v) { return {gsign(v.x), gsign(v.y), gsign(v.z)}; }
struct vec4 {
    float r, g, b, a;
};

struct uvec2 {
    uint32_t r, g;
};

struct Ctx {
    // uniforms
    vec3 CameraPosition, CameraUp, CameraLeft, VolExtentMin, VolExtentMax;
    vec3 BrickStepVector, BrickDimension, BrickPoolDimension;
    float RayStepSize, RootVoxelHalfSize, PixelSize, LodScalar;
    int OctTreeDepthMinusOne, MaxNodesToPushThisFrame, MaxTraversals;
    bool RootNodeIsConstant;
    // ray generation
    vec3 look, right, upv, volScale;
    int W, H;
    // textures
    const uint32_t* nodes;
    uint32_t ndim[3];
    const uint32_t* pool;
    uint32_t pdim[3];
    const float* jitter;
    // wide pool (vox_brick_pool_create_wide): word 1 of a resident node is
    // its slot; the brick is sampled in its own (B+2)^3 texel frame
    bool wide = false;
    uint32_t wideSlots = 0;
};

uvec2 texelFetch(const Ctx& c, uint32_t x, uint32_t y, uint32_t z) {
    if (x >= c.ndim[0] || y >= c.ndim[1] || z >= c.ndim[2]) return {0, 0};
    const uint64_t t = ((uint64_t)z * c.ndim[1] + y) * c.ndim[0] + x;
    return {c.nodes[2 * t], c.nodes[2 * t + 1]};
}

vec4 unorm(uint32_t v) {
    return {(float)(v & 255u) / 255.0f, (float)((v >> 8) & 255u) / 255.0f,
            (float)((v >> 16) & 255u) / 255.0f, (float)(v >> 24) / 255.0f};
}

// texture(BrickSampler, s)
vec4 texture(const Ctx& c, vec3 s) {
    int i0[3];
    float w[3];
    for (int a = 0; a < 3; ++a) {
        float u = s[a];
        if (std::isnan(u)) u = 0.0f;
        const float b = u * c.BrickPoolDimension[a] - 0.5f;
        const float f = std::floor(b);
        w[a] = std::rint((b - f) * 256.0f) / 256.0f;
        i0[a] = (int)std::fmax(std::fmin(f, 1.0e9f), -1.0e9f);
    }
    vec4 r = {0, 0, 0, 0};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                int ix = i0[0] + dx, iy = i0[1] + dy, iz = i0[2] + dz;
                ix = std::min(std::max(ix, 0), (int)c.pdim[0] - 1);
                iy = std::min(std::max(iy, 0), (int)c.pdim[1] - 1);
                iz = std::min(std::max(iz, 0), (int)c.pdim[2] - 1);
                const float wk = ((dx ? w[0] : 1.0f - w[0]) * (dy ? w[1] : 1.0f - w[1])) *
                                 (dz ? w[2] : 1.0f - w[2]);
                const vec4 t = unorm(c.pool[((uint64_t)iz * c.pdim[1] + iy) * c.pdim[0] + ix]);
                r.r = r.r + wk * t.r;
                r.g = r.g + wk * t.g;
                r.b = r.b + wk * t.b;
                r.a = r.a + wk * t.a;
            }
    return r;
}

float gpow2n(float x, int n) {
    for (int i = 0; i < n; ++i) x = x * x;
    return x;
}

// NodePlanes (GigaVoxels.frag 46-60) and ComputeMinRayPlaneIntersection /
// ComputeMinRayBoxIntersection (131-180).
struct plane4 {
    float x, y, z, w;
};
void ComputeMinRayPlaneIntersection(vec3 rayPos, vec3 rayDir, vec3 pointOnPlane, plane4 plane,
                                    float& tMin) {
    const vec3 n(plane.x, plane.y, plane.z);
    float dirDotN = dot(rayDir, n);
    if (dirDotN > 0) {
        vec3 v = rayPos - pointOnPlane;
        float tIsect = -(dot(n, v)) / dirDotN;
        if (tIsect < tMin) {
            if (std::fabs(tIsect) > 0.0f) tMin = tIsect;
        }
    }
}
float ComputeMinRayBoxIntersection(vec3 rayPos, vec3 rayDir, vec3 mn, vec3 mx) {
    plane4 P[6] = {{-1, 0, 0, mn.x}, {0, -1, 0, mn.y}, {0, 0, -1, mn.z},
                   {1, 0, 0, -mx.x}, {0, 1, 0, -mx.y}, {0, 0, 1, -mx.z}};
    float tMin = 4.0f;
    for (int i = 0; i < 6; ++i) ComputeMinRayPlaneIntersection(rayPos, rayDir, i < 3 ? mn : mx, P[i], tMin);
    return tMin;
}

void AdvanceRay(vec3& pos, vec3 dir, float dist, float step) {
    int numSteps = (int)(dist / step) + 1;
    pos = pos + (dir * ((float)numSteps * step));
}

void blend(vec4& dst, vec4 src) {
    const float k = 1.0f - dst.a;
    dst.r = k * src.r + dst.r;
    dst.g = k * src.g + dst.g;
    dst.b = k * src.b + dst.b;
    dst.a = k * src.a + dst.a;
}

void TraverseConstant(vec4 constValue, vec3& pos, vec3 rayDir, float rayStepSize, float sampleRatio,
                      int ratioLog2, vec3 mn, vec3 mx, vec4& dst) {
    float timeToExit = ComputeMinRayBoxIntersection(pos, rayDir, mn, mx);
    if (constValue.a < 0.005f) {
        AdvanceRay(pos, rayDir, timeToExit, rayStepSize);
        return;
    }
    int numSteps = (int)(timeToExit / rayStepSize) + 1;
    vec3 worldRayStep = rayDir * rayStepSize;
    for (int i = 0; i < numSteps; ++i) {
        pos = pos + worldRayStep;
        if (dst.a > 0.995f) {
            numSteps -= (i + 1);
            pos = pos + (rayDir * ((float)numSteps * rayStepSize));
            return;
        }
        vec4 src = constValue;
        src.a = 1.0f - gpow2n(1.0f - src.a, ratioLog2);
        src.r *= sampleRatio;
        src.g *= sampleRatio;
        src.b *= sampleRatio;
        blend(dst, src);
    }
}

void TraverseBrick(const Ctx& c, vec3& pos, vec3 rayDir, vec3 worldRayStep, vec3 curBrickRayPos,
                   vec3 brickRayStep, float sampleRatio, vec3 mn, vec3 mx, vec4& dst) {
    float timeToExit = ComputeMinRayBoxIntersection(pos, rayDir, mn, mx);
    float rayStepSize = c.RayStepSize * sampleRatio;
    int numSteps = (int)(timeToExit / rayStepSize) + 1;
    for (int i = 0; i < numSteps; ++i) {
        pos = pos + worldRayStep;
        if (dst.a > 0.995f) {
            numSteps -= (i + 1);
            pos = pos + (rayDir * ((float)numSteps * rayStepSize));
            return;
        }
        vec4 src = texture(c, curBrickRayPos);
        src.a = 1.0f - (1.0f - src.a);  // pow(1 - a, 1.0)
        blend(dst, src);
        curBrickRayPos = curBrickRayPos + brickRayStep;
    }
}

float ComputeVoxelSize(const Ctx& c, vec3 rayPosition, float h) {
    vec3 camToLL = normalize(((rayPosition - (c.CameraUp * h)) + (c.CameraLeft * h)) - c.CameraPosition);
    vec3 camToUR = normalize(((rayPosition + (c.CameraUp * h)) - (c.CameraLeft * h)) - c.CameraPosition);
    return 1.0f - dot(camToLL, camToUR);
}

uint32_t uclamp01(float v) { return v >= 1.0f ? 1u : 0u; }  // clamp(uint(v), 0, 1), v < 0 -> 0

// main (GigaVoxels.frag 544-805) for pixel (px, py).
void shade(const Ctx& c, int px, int py, uint32_t usage[12], uint32_t& rgba) {
    for (int k = 0; k < 12; ++k) usage[k] = 0;
    vec4 dst = {0, 0, 0, 0};
    // the rasterized entry point
    const float sx = (2.0f * ((float)px + 0.5f)) / (float)c.W - 1.0f;
    const float sy = (2.0f * ((float)py + 0.5f)) / (float)c.H - 1.0f;
    const vec3 dw = (c.look + (c.right * sx)) + (c.upv * sy);
    const vec3 dn = dw / c.volScale;
    float tEnter = -INFINITY, tExit = INFINITY;
    bool hit = true;
    for (int a = 0; a < 3; ++a) {
        if (dn[a] == 0.0f) {
            if (c.CameraPosition[a] < c.VolExtentMin[a] || c.CameraPosition[a] > c.VolExtentMax[a])
                hit = false;
            continue;
        }
        const float t0 = (c.VolExtentMin[a] - c.CameraPosition[a]) / dn[a];
        const float t1 = (c.VolExtentMax[a] - c.CameraPosition[a]) / dn[a];
        tEnter = std::fmax(tEnter, std::fmin(t0, t1));
        tExit = std::fmin(tExit, std::fmax(t0, t1));
    }
    const float tStart = std::fmax(tEnter, 0.0f);
    if (!(tExit >= tStart)) hit = false;
    if (hit) {
        vec3 rayWorldPosition = c.CameraPosition + (dn * tStart);
        vec3 rayDir = normalize(dn);
        float jitterOffset = c.jitter[(py % 32) * 32 + (px % 32)];
        rayWorldPosition = rayWorldPosition + ((rayDir * c.RayStepSize) * jitterOffset);
        vec3 test1 = sign(rayWorldPosition - c.VolExtentMin);
        vec3 test2 = sign(c.VolExtentMax - rayWorldPosition);
        float inside = dot(test1, test2);
        int traversalCount = 0;
        if (inside >= 3.0f) {
            int curListIndex = 0, curComponentIndex = 0;
            uint32_t nuX = 0, nuY = 0, nuZ = 0;
            int pcx = px % 2, pcy = py % 2;
            int startPushingNodes = (pcx * 12) + (pcy * 12 * 2);
            int stopPushingNodes = startPushingNodes + 12;
            int maxNodesPerPixelBlock = 12 * 2 * 2;
            vec3 brickRayStep = rayDir * c.BrickStepVector;
            uint32_t ptr[3] = {0, 0, 0};
            uvec2 rootNode = texelFetch(c, 0, 0, 0);
            uvec2 node = rootNode;
            vec3 nodeExtentMin = c.VolExtentMin, nodeExtentMax = c.VolExtentMax;
            float nodeSize = 1.0f;
            float voxelHalfSize = c.RootVoxelHalfSize;
            vec3 rayBrickPosition = rayWorldPosition;
            vec3 curBrickRayPosition = rayWorldPosition;
            uvec2 curBrickPtr = rootNode;
            bool brickIsLoaded = false, curNodeIsConstant = c.RootNodeIsConstant;
            float projVoxelSize = 0.0f;
            int curDepth = 0, curBrickDepth = 0;
            uint32_t child[3] = {0, 0, 0};
            int nodeListTraversalCount = 0, pushedNodeCount = 0;
            while (inside >= 3.0f && traversalCount < c.MaxTraversals) {
                curNodeIsConstant = (node.r & 0x40000000u) != 0u;
                brickIsLoaded = (node.r & 0x80000000u) != 0u;
                if (brickIsLoaded) {
                    curBrickPtr = node;
                    curBrickRayPosition = rayBrickPosition;
                    curBrickDepth = curDepth;
                }
                projVoxelSize = ComputeVoxelSize(c, rayWorldPosition, voxelHalfSize) * c.LodScalar;
                bool terminal = (!curNodeIsConstant && !brickIsLoaded) || c.PixelSize > projVoxelSize;
                if (!terminal) {
                    child[0] = (node.r >> 20u) & 0x3FFu;
                    child[1] = (node.r >> 10u) & 0x3FFu;
                    child[2] = node.r & 0x3FFu;
                    terminal = (child[0] == 0 && child[1] == 0 && child[2] == 0) ||
                               curDepth >= c.OctTreeDepthMinusOne;  // bound for damaged data
                }
                if (terminal) {
                    nuX = ptr[0];
                    nuY = ptr[1];
                    nuZ = ptr[2];
                    if (curNodeIsConstant) {
                        const int n = c.OctTreeDepthMinusOne - curDepth;
                        float sampleRatio = std::ldexp(1.0f, n);
                        vec4 cv = unorm(node.g);
                        TraverseConstant(cv, rayWorldPosition, rayDir, c.RayStepSize * sampleRatio,
                                         sampleRatio, n, nodeExtentMin, nodeExtentMax, dst);
                    } else {
                        vec3 bp((float)((curBrickPtr.g >> 20u) & 0x3FFu) / c.BrickPoolDimension.x,
                                (float)((curBrickPtr.g >> 10u) & 0x3FFu) / c.BrickPoolDimension.y,
                                (float)(curBrickPtr.g & 0x3FFu) / c.BrickPoolDimension.z);
                        Ctx cb = c;
                        if (c.wide) {
                            // the slot's first interior texel in the brick's own frame;
                            // a word 1 that is no slot (no resident brick on the
                            // path) reads slot 0
                            bp = vec3(1.0f / c.BrickPoolDimension.x, 1.0f / c.BrickPoolDimension.y,
                                      1.0f / c.BrickPoolDimension.z);
                            const uint32_t slot = curBrickPtr.g < c.wideSlots ? curBrickPtr.g : 0u;
                            cb.pool = c.pool + (uint64_t)slot * c.pdim[0] * c.pdim[1] * c.pdim[2];
                        }
                        vec3 curBrickRayPos =
                            bp + ((curBrickRayPosition * c.BrickDimension) / c.BrickPoolDimension);
                        float sampleRatio = std::ldexp(1.0f, c.OctTreeDepthMinusOne - curBrickDepth);
                        vec3 worldRayStep = rayDir;
                        worldRayStep = worldRayStep * c.RayStepSize;
                        worldRayStep = worldRayStep * sampleRatio;
                        TraverseBrick(cb, rayWorldPosition, rayDir, worldRayStep, curBrickRayPos,
                                      brickRayStep, sampleRatio, nodeExtentMin, nodeExtentMax, dst);
                    }
                    if (nodeListTraversalCount >= startPushingNodes &&
                        nodeListTraversalCount < stopPushingNodes &&
                        pushedNodeCount < c.MaxNodesToPushThisFrame) {
                        ++pushedNodeCount;
                        // AddToNodeUsageList (508-542)
                        uint32_t bits = 0u;
                        bits |= ((nuX + 1u) << 20u);
                        bits |= ((nuY + 1u) << 10u);
                        bits |= (nuZ + 1u);
                        usage[curListIndex * 4 + curComponentIndex] = bits;
                        ++curComponentIndex;
                        if (curComponentIndex == 4) {
                            curComponentIndex = 0;
                            ++curListIndex;
                            if (curListIndex == 3) curListIndex = 0;
                        }
                    }
                    if (dst.a > 0.995f && (curNodeIsConstant || brickIsLoaded)) break;
                    node = rootNode;  // octTreeNodePtr keeps its value
                    nodeExtentMin = c.VolExtentMin;
                    nodeExtentMax = c.VolExtentMax;
                    nodeSize = 1.0f;
                    voxelHalfSize = c.RootVoxelHalfSize;
                    rayBrickPosition = rayWorldPosition;
                    curBrickRayPosition = rayWorldPosition;
                    curBrickPtr = node;
                    brickIsLoaded = false;
                    curNodeIsConstant = c.RootNodeIsConstant;
                    projVoxelSize = 0.0f;
                    curDepth = 0;
                    curBrickDepth = 0;
                    child[0] = child[1] = child[2] = 0;
                    test1 = sign(rayWorldPosition - c.VolExtentMin);
                    test2 = sign(c.VolExtentMax - rayWorldPosition);
                    inside = dot(test1, test2);
                    ++traversalCount;
                    ++nodeListTraversalCount;
                    if (nodeListTraversalCount >= maxNodesPerPixelBlock) nodeListTraversalCount = 0;
                } else {
                    vec3 brickOffset = rayBrickPosition * 2.0f;
                    uint32_t co[3] = {uclamp01(brickOffset.x), uclamp01(brickOffset.y),
                                      uclamp01(brickOffset.z)};
                    rayBrickPosition = brickOffset - vec3((float)co[0], (float)co[1], (float)co[2]);
                    for (int a = 0; a < 3; ++a) ptr[a] = child[a] + co[a];
                    ++curDepth;
                    nodeSize *= 0.5f;
                    nodeExtentMin = nodeExtentMin + (vec3((float)co[0], (float)co[1], (float)co[2]) * nodeSize);
                    nodeExtentMax = nodeExtentMax - ((vec3(1.0f) - vec3((float)co[0], (float)co[1], (float)co[2])) * nodeSize);
                    voxelHalfSize *= 0.5f;
                    node = texelFetch(c, ptr[0], ptr[1], ptr[2]);
                }
            }
        }
    }
    float v[4] = {dst.r, dst.g, dst.b, dst.a};
    rgba = 0;
    for (int k = 0; k < 4; ++k) {
        float x = v[k] * 255.0f;
        x = x > 255.0f ? 255.0f : (x > 0.0f ? x : 0.0f);
        rgba |= (uint32_t)x << (8 * k);
    }
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" int32_t ref_raycast_pool(const ref_ray_camera* cam, const ref_ray_params* prm,
                                    int32_t levels, const uint32_t B[3], const uint32_t* nodes,
                                    const uint32_t ndim[3], const uint32_t* pool,
                                    const uint32_t pdim[3], int32_t wide, uint32_t* usage_out,
                                    uint32_t* rgba_out) {
    Ctx c;
    c.wide = wide != 0;
    // a wide pool: slots stacked along z, each sampled in its own frame
    const uint32_t fdim[3] = {wide ? B[0] + 2 : pdim[0], wide ? B[1] + 2 : pdim[1],
                              wide ? B[2] + 2 : pdim[2]};
    c.wideSlots = wide ? pdim[2] / (B[2] + 2) : 0u;
    // GigaVoxelsRenderer::drawVolume uniforms (1650-1816)
    float T[3], S[3];
    for (int a = 0; a < 3; ++a) {
        T[a] = (prm->vol_min[a] + prm->vol_max[a]) * 0.5f;
        S[a] = prm->vol_max[a] - prm->vol_min[a];
        if (!(S[a] > 0.0f)) return 0;
        c.volScale[a] = S[a];
        c.VolExtentMin[a] = (prm->vol_min[a] - T[a]) / S[a] + 0.5f;
        c.VolExtentMax[a] = (prm->vol_max[a] - T[a]) / S[a] + 0.5f;
        c.CameraPosition[a] = (cam->position[a] - T[a]) / S[a] + 0.5f;
        c.CameraUp[a] = cam->up[a];
    }
    c.CameraLeft = vec3(cam->up[1] * cam->look[2] - cam->up[2] * cam->look[1],
                        cam->up[2] * cam->look[0] - cam->up[0] * cam->look[2],
                        cam->up[0] * cam->look[1] - cam->up[1] * cam->look[0]);
    const float tanY = std::tan(cam->fov_y_deg * 0.5f * 3.14159265358979f / 180.0f);
    const float tanX = tanY * (float)cam->width / (float)cam->height;
    for (int a = 0; a < 3; ++a) {
        c.look[a] = cam->look[a];
        c.right[a] = -c.CameraLeft[a] * tanX;
        c.upv[a] = cam->up[a] * tanY;
    }
    const uint32_t Nmax = std::max(B[0], std::max(B[1], B[2])) << (levels - 1);
    c.RayStepSize = prm->ray_step > 0.0f ? prm->ray_step : (1.0f / (float)Nmax) * 0.5f;
    if (!(c.RayStepSize >= 1.0e-5f)) return 0;
    c.LodScalar = prm->lod_scalar > 0.0f ? prm->lod_scalar : 10.0f;
    const float pow2 = std::ldexp(1.0f, levels - 1);
    for (int a = 0; a < 3; ++a) {
        const float rayStepRoot = (S[a] * c.RayStepSize) * pow2;
        const float rootVoxel = S[a] / (float)B[a];
        const float ratio = rayStepRoot / rootVoxel;
        c.BrickStepVector[a] = (1.0f / (float)fdim[a]) * ratio;
        c.BrickDimension[a] = (float)B[a];
        c.BrickPoolDimension[a] = (float)fdim[a];
        c.ndim[a] = ndim[a];
        c.pdim[a] = fdim[a];
    }
    const float bx = (float)B[0], by = (float)B[1], bz = (float)B[2];
    const float minB = bx < ((by < bz) ? by : bz) ? bx : bz;
    c.RootVoxelHalfSize = (1.0f / minB) * 0.5f;
    {
        const double W = cam->width, H = cam->height;
        double r0[3], r1[3];
        for (int a = 0; a < 3; ++a) {
            const double sx0 = 2.0 * (W * 0.5) / W - 1.0, sy0 = 2.0 * (H * 0.5) / H - 1.0;
            const double sx1 = 2.0 * (W * 0.5 + 1.0) / W - 1.0, sy1 = 2.0 * (H * 0.5 + 1.0) / H - 1.0;
            r0[a] = c.look[a] + (double)c.right[a] * sx0 + (double)c.upv[a] * sy0;
            r1[a] = c.look[a] + (double)c.right[a] * sx1 + (double)c.upv[a] * sy1;
        }
        const double n0 = std::sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
        const double n1 = std::sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
        const double cosA = (r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2]) / (n0 * n1);
        c.PixelSize = (float)(1.0 - cosA);
    }
    c.OctTreeDepthMinusOne = levels - 1;
    c.MaxNodesToPushThisFrame = 12 * (int)((cam->frame % 2) + 1);
    c.MaxTraversals = prm->max_traversals ? (int)prm->max_traversals : 4096;
    c.W = cam->width;
    c.H = cam->height;
    c.nodes = nodes;
    c.pool = pool;
    c.RootNodeIsConstant = (nodes[0] & 0x40000000u) != 0u;
    std::vector<float> jit(1024);
    uint64_t s = prm->jitter_seed;
    for (int k = 0; k < 1024; ++k) jit[k] = (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f);
    c.jitter = jit.data();
    for (int py = 0; py < c.H; ++py)
        for (int px = 0; px < c.W; ++px) {
            const uint64_t pix = (uint64_t)py * c.W + px;
            uint32_t rgba;
            shade(c, px, py, usage_out + 12 * pix, rgba);
            if (rgba_out) rgba_out[pix] = rgba;
        }
    return 1;
}

extern "C" int32_t ref_raycast(const ref_ray_camera* cam, const ref_ray_params* prm, int32_t levels,
                               const uint32_t B[3], const uint32_t* nodes, const uint32_t ndim[3],
                               const uint32_t* pool, const uint32_t pdim[3], uint32_t* usage_out,
                               uint32_t* rgba_out) {
    return ref_raycast_pool(cam, prm, levels, B, nodes, ndim, pool, pdim, 0, usage_out, rgba_out);
}
