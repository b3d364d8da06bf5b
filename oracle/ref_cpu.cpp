// ref_cpu.cpp -- CPU restatement of the reference voxelizer and brick writer.
//
// TEST INFRASTRUCTURE ONLY: the oracle for the HIP path.  Loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
// product library.  Build: oracle/Makefile (g++ -O2 -ffp-contract=off).
//
// All citations are to apecoraro/VoxVizThesis, VoxViz/Voxelizer/Voxelizer/:
//   VK = VoxelizerKernels.cu, VZ = Voxelizer.cu, VH = Voxelizer.h,
//   BW = VoxelBrickWriter.cu.
//
// Scope: the single-chunk regime, i.e. a chunk that holds the whole voxel
// space S = N + 4*extra (extra = (N/B)>>1, VZ:142-144).  In that regime every
// brick at every level lies strictly inside the chunk, so the partial-brick
// assembly of BW:1026-1210 never runs and chunk-seam typing (VZ:1154-1172)
// cannot arise.  A smaller chunk is rejected, exactly as the HIP path does.
//
// Semantics the reference leaves open are fixed as follows (same choices in
// the HIP path):
//   * slot race (VK:507-514): triangles are applied in ascending index order,
//     so slots 0..2 hold the three lowest hitting indices and slot 3 the
//     highest once four or more triangles hit.
//   * terrain fill race (VK:888-917): every thread reads its voxel first;
//     voxels with new hits keep their own value; other voxels take the value
//     of the nearest voxel above (<= 19) that has new hits in this group.
//   * triangle/vertex indices are 32-bit (the reference narrows them to
//     short at VK:775-786, which wraps past 10,922 triangles).
//   * float -> uchar conversion saturates to [0,255] and maps NaN to 0 (the
//     reference's static_cast is undefined outside that range).
//   * bilinear texture fetch: CUDA's documented linear filter with the
//     fractional weights held in 8 fractional bits (round to nearest); NaN
//     coordinates read as 0.
#include "ref_cpu.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>


namespace {

std::string g_error;

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

inline float comp(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divv(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// GLM 0.9.3 dot: left-to-right sum of products.
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
// GLM 0.9.3 cross component order.
inline V3 cross3(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// GLM 0.9.3 normalize: v * inversesqrt(dot(v,v)), inversesqrt = 1/sqrt.
inline V3 normalize3(V3 v) {
    float sqr = v.x * v.x + v.y * v.y + v.z * v.z;
    float inv = 1.0f / std::sqrt(sqr);
    return muls(v, inv);
}
inline float gmin(float x, float y) { return x < y ? x : y; }
inline float gmax(float x, float y) { return x > y ? x : y; }
inline float gclamp(float x, float lo, float hi) { return gmax(gmin(x, hi), lo); }

// float -> int truncation, saturating, NaN -> 0 (reference: glm::ivec3(float)).
inline int64_t trunc_i(float f) {
    if (!(f == f)) return 0;
    if (f >= 9.0e18f) return INT64_C(9000000000000000000);
    if (f <= -9.0e18f) return -INT64_C(9000000000000000000);
    return (int64_t)f;
}
// float -> unsigned char, truncating (VK:681-687, BW:563-568), saturating.
inline uint8_t quant_u8(float f) {
    if (!(f == f)) return 0;
    if (f <= 0.0f) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)(int)f;
}
inline uint32_t pack4(uint8_t r, uint8_t g, uint8_t b, uint8_t a) {
    return (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24);
}
// VoxColorToVec4 (VK:657-663).
inline V4 unpack4(uint32_t c) {
    return {(float)(c & 255u) / 255.0f, (float)((c >> 8) & 255u) / 255.0f,
            (float)((c >> 16) & 255u) / 255.0f, (float)(c >> 24) / 255.0f};
}
// Vec4ToVoxColor (VK:680-687).
inline uint32_t vec4_to_color(V4 c) {
    return pack4(quant_u8(c.x * 255.0f), quant_u8(c.y * 255.0f), quant_u8(c.z * 255.0f),
                 quant_u8(c.w * 255.0f));
}

struct Tri {
    V3 v[3];
    V3 e[3];
    V3 n;
    uint32_t bmin[3], bmax[3];
};

// TrianglePlaneOverlapsVoxel (VK:17-53).
bool plane_overlaps(const V3& v0, const V3& n, const V3& p, const V3& dp, const V3& uvw) {
    V3 voxOffset = mulv(uvw, dp);
    V3 p0 = add(p, voxOffset);
    const V3 cd[7] = {
        dp,
        {dp.x, dp.y, 0.0f},
        {dp.x, 0.0f, 0.0f},
        {0.0f, 0.0f, dp.z},
        {0.0f, dp.y, dp.z},
        {dp.x, 0.0f, dp.z},
        {0.0f, dp.y, 0.0f},
    };
    float D = dot3(neg(n), v0);
    float dist1 = dot3(n, p0) + D;
    for (int i = 0; i < 7; ++i) {
        V3 p1 = add(p0, cd[i]);
        float dist2 = dot3(n, p1) + D;
        if (dist1 * dist2 <= 0.0f) return true;
    }
    return false;
}

// TriangleOverlapsVoxelOnAxis (VK:93-144).
bool axis_overlaps(const Tri& t, const V3& p, const V3& dp, const V3& uvw, int a1, int a2,
                   int other) {
    float projScalar = comp(t.n, other) >= 0.0f ? 1.0f : -1.0f;
    float ot[3];
    for (int i = 0; i < 3; ++i) {
        V2 en = {(-comp(t.e[i], a2)) * projScalar, comp(t.e[i], a1) * projScalar};
        V2 vp = {en.x * comp(dp, a1), en.y * comp(dp, a2)};
        V2 nen = {-en.x, -en.y};
        float dpp = dot2(nen, V2{comp(t.v[i], a1), comp(t.v[i], a2)}) +
                    dot2(en, V2{comp(p, a1), comp(p, a2)}) +
                    gmax(0.0f, comp(dp, a1) * en.x) + gmax(0.0f, comp(dp, a2) * en.y);
        ot[i] = dpp + comp(uvw, a2) * vp.y + comp(uvw, a1) * vp.x;
    }
    return ot[0] >= 0.0f && ot[1] >= 0.0f && ot[2] >= 0.0f;
}

// TriangleOverlapsVoxel (VK:146-170): XY, ZX, YZ.
bool tri_overlaps(const Tri& t, const V3& p, const V3& dp, const V3& uvw) {
    if (!axis_overlaps(t, p, dp, uvw, 0, 1, 2)) return false;
    if (!axis_overlaps(t, p, dp, uvw, 2, 0, 1)) return false;
    return axis_overlaps(t, p, dp, uvw, 1, 2, 0);
}

// ComputeVoxelBarycentricCoords (VK:602-624).
V2 bary(V2 P, V2 A, V2 B, V2 C) {
    V2 v0 = {C.x - A.x, C.y - A.y};
    V2 v1 = {B.x - A.x, B.y - A.y};
    V2 v2 = {P.x - A.x, P.y - A.y};
    float dot00 = dot2(v0, v0);
    float dot01 = dot2(v0, v1);
    float dot02 = dot2(v0, v2);
    float dot11 = dot2(v1, v1);
    float dot12 = dot2(v1, v2);
    float invDenom = 1.0f / (dot00 * dot11 - dot01 * dot01);
    return {gclamp((dot11 * dot02 - dot01 * dot12) * invDenom, 0.0f, 1.0f),
            gclamp((dot00 * dot12 - dot01 * dot02) * invDenom, 0.0f, 1.0f)};
}
// ComputeVoxelTexCoords (VK:626-640).
V2 interp_uv(V2 uv, V2 a, V2 b, V2 c) {
    V2 V = {(b.x - a.x) * uv.y, (b.y - a.y) * uv.y};
    V2 U = {(c.x - a.x) * uv.x, (c.y - a.y) * uv.x};
    return {a.x + U.x + V.x, a.y + U.y + V.y};
}
// ComputeVoxelNormal (VK:642-655).
V3 interp_n(V2 uv, V3 a, V3 b, V3 c) {
    V3 V = muls(sub(b, a), uv.y);
    V3 U = muls(sub(c, a), uv.x);
    return add(add(a, U), V);
}

struct Image {
    const uint8_t* px;
    int w, h, addr0, addr1;
};

int wrap_index(int i, int n, int mode, bool* border) {
    switch (mode) {
        case 0: {  // wrap
            int m = i % n;
            return m < 0 ? m + n : m;
        }
        case 2: {  // mirror
            int period = 2 * n;
            int m = i % period;
            if (m < 0) m += period;
            return m < n ? m : period - 1 - m;
        }
        case 3:  // border
            if (i < 0 || i >= n) { *border = true; return 0; }
            return i;
        default:  // clamp
            return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
}

V4 texel(const Image& im, int i, int j) {
    bool border = false;
    int x = wrap_index(i, im.w, im.addr0, &border);
    int y = wrap_index(j, im.h, im.addr1, &border);
    if (border) return {0.0f, 0.0f, 0.0f, 0.0f};
    const uint8_t* q = im.px + 4 * ((size_t)y * (size_t)im.w + (size_t)x);
    return {(float)q[0] / 255.0f, (float)q[1] / 255.0f, (float)q[2] / 255.0f,
            (float)q[3] / 255.0f};
}

// tex2D with cudaFilterModeLinear and normalized coordinates (VK:586-600,
// 835-837): T = (1-a)(1-b)T00 + a(1-b)T10 + (1-a)bT01 + abT11 with a, b held
// in 8 fractional bits.
V4 tex2d(const Image& im, float u, float v) {
    if (!(u == u)) u = 0.0f;
    if (!(v == v)) v = 0.0f;
    float xb = u * (float)im.w - 0.5f;
    float yb = v * (float)im.h - 0.5f;
    float fx = std::floor(xb), fy = std::floor(yb);
    float ax = std::nearbyint((xb - fx) * 256.0f) / 256.0f;
    float ay = std::nearbyint((yb - fy) * 256.0f) / 256.0f;
    int i = (int)std::max(std::min(fx, 1.0e9f), -1.0e9f);
    int j = (int)std::max(std::min(fy, 1.0e9f), -1.0e9f);
    V4 t00 = texel(im, i, j), t10 = texel(im, i + 1, j);
    V4 t01 = texel(im, i, j + 1), t11 = texel(im, i + 1, j + 1);
    float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay);
    float w01 = (1.0f - ax) * ay, w11 = ax * ay;
    V4 r;
    r.x = ((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x;
    r.y = ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y;
    r.z = ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z;
    r.w = ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w;
    return r;
}

struct Level {
    uint32_t d[3];
    std::vector<uint32_t> col;
    std::vector<float> nrm;  // 3 per voxel
    size_t idx(size_t x, size_t y, size_t z) const { return (z * d[1] + y) * d[0] + x; }
};

struct TreeLevel {
    std::vector<uint8_t> types;
    std::vector<uint32_t> cc;
    std::vector<uint32_t> bidx;
    std::vector<uint32_t> bcol, bnrm;
    uint64_t nb = 0;
};

}  // namespace

struct ref_result {
    uint32_t N[3], B[3], S[3], extra[3];
    int L = 0;
    std::vector<uint32_t> occ;
    std::vector<Level> lv;        // mip level k (0 = full res)
    std::vector<TreeLevel> tree;  // tree depth d (0 = root)
    std::vector<double> n64;      // fp64 twin normals at level 0 (optional)
    bool has64 = false;
};

extern "C" const char* ref_error(void) { return g_error.c_str(); }

extern "C" ref_result* ref_run(const uint32_t N[3], const uint32_t B[3], const float pin[3],
                               const float delta[3], const ref_group* groups, int32_t ngroups,
                               int32_t flags, int32_t* status) {
    *status = 0;
    g_error.clear();
    ref_result* r = new ref_result();
    // Voxelizer ctor (VZ:125-153): extra = (N/B)>>1; levels by doubling B.
    int L = 0;
    for (int a = 0; a < 3; ++a) {
        if (B[a] < 2 || N[a] % B[a] != 0) { g_error = "N must be a multiple of B >= 2"; delete r; return nullptr; }
        uint32_t ratio = N[a] / B[a];
        if (ratio < 2 || (ratio & (ratio - 1)) != 0) { g_error = "N/B must be a power of two >= 2"; delete r; return nullptr; }
        int l = 1;
        while ((B[a] << (l - 1)) != N[a]) ++l;
        if (a > 0 && l != L) { g_error = "N/B must be equal on all axes"; delete r; return nullptr; }
        L = l;
        r->N[a] = N[a];
        r->B[a] = B[a];
        r->extra[a] = ratio >> 1;
        r->S[a] = N[a] + 4 * r->extra[a];
    }
    r->L = L;
    const uint32_t* S = r->S;
    const V3 dP = {delta[0], delta[1], delta[2]};
    const V3 P = {pin[0], pin[1], pin[2]};
    // setVoxelizationParams (VZ:155-161): offsetP = p - extra*delta.
    const V3 offP = sub(P, mulv(V3{(float)r->extra[0], (float)r->extra[1], (float)r->extra[2]}, dP));
    const size_t nvox = (size_t)S[0] * S[1] * S[2];
    const bool twin = (flags & REF_FP64_TWIN) != 0;

    // Concatenate groups (allocateTriangleMemory, VZ:346-457).
    std::vector<Tri> tris;
    std::vector<size_t> grpOff(ngroups + 1, 0);
    for (int g = 0; g < ngroups; ++g) grpOff[g + 1] = grpOff[g] + groups[g].ntri;
    tris.resize(grpOff[ngroups]);
    // ComputeEdgesFaceNormalsAndBounds (VK:230-347) with voxDim = S (VZ:472).
    for (int g = 0; g < ngroups; ++g) {
        for (uint64_t t = 0; t < groups[g].ntri; ++t) {
            Tri& T = tris[grpOff[g] + t];
            const float* q = groups[g].verts + 9 * t;
            for (int k = 0; k < 3; ++k) T.v[k] = {q[3 * k], q[3 * k + 1], q[3 * k + 2]};
            T.e[0] = sub(T.v[1], T.v[0]);
            T.e[1] = sub(T.v[2], T.v[1]);
            T.e[2] = sub(T.v[0], T.v[2]);
            V3 p0p2 = sub(T.v[2], T.v[0]);
            T.n = normalize3(cross3(T.e[0], p0p2));
            for (int k = 0; k < 3; ++k) {
                V3 diff = divv(sub(T.v[k], offP), dP);
                for (int a = 0; a < 3; ++a) {
                    int64_t iv = trunc_i(comp(diff, a));
                    int64_t hi = (int64_t)S[a] - 1;
                    uint32_t u = (uint32_t)(iv < 0 ? 0 : (iv > hi ? hi : iv));
                    if (k == 0) { T.bmin[a] = u; T.bmax[a] = u; }
                    else if (u < T.bmin[a]) T.bmin[a] = u;
                    else if (u > T.bmax[a]) T.bmax[a] = u;
                }
            }
        }
    }

    // Per-voxel chunk state (allocateDeviceVoxelizationMemory VZ:1349-1411,
    // memsets VZ:2325-2332).
    std::vector<uint32_t> counts(nvox, 0);
    std::vector<int32_t> slots(4 * nvox, -1);
    std::vector<uint32_t> colors(nvox, 0);
    std::vector<float> normals(3 * nvox, 0.0f);
    std::vector<uint8_t> newHit(nvox, 0);
    r->occ.assign((nvox + 31) / 32, 0u);
    if (twin) { r->n64.assign(3 * nvox, 0.0); r->has64 = true; }
    auto VI = [&](size_t x, size_t y, size_t z) { return (z * S[1] + y) * S[0] + x; };

    bool anyOverlap = false;
    for (int g = 0; g < ngroups; ++g) {
        const ref_group& G = groups[g];
        // Group voxel bbox (addTriangleGroup, VH:137-144): clamp to [0, S], cast.
        uint32_t gmn[3], gmx[3];
        for (int a = 0; a < 3; ++a) {
            float lo = gclamp((G.bbox_min[a] - comp(offP, a)) / comp(dP, a), 0.0f, (float)S[a]);
            float hi = gclamp((G.bbox_max[a] - comp(offP, a)) / comp(dP, a), 0.0f, (float)S[a]);
            gmn[a] = (uint32_t)trunc_i(lo);
            gmx[a] = (uint32_t)trunc_i(hi);
        }
        // Chunk overlap test (VZ:2341-2353); chunk = [0, S-1].
        bool overlaps = true;
        for (int a = 0; a < 3; ++a)
            if (gmn[a] > S[a] - 1) overlaps = false;
        if (!overlaps) continue;
        anyOverlap = true;
        uint32_t zStart = gmn[2], zEnd = std::min(gmx[2], S[2] - 1);

        // ComputeVoxelization (VK:349-530), one triangle at a time in index
        // order (the cpuTest loop shape, VK:2426-2505).
        for (size_t t = grpOff[g]; t < grpOff[g + 1]; ++t) {
            const Tri& T = tris[t];
            uint32_t z0 = std::max(T.bmin[2], zStart), z1 = std::min(T.bmax[2], zEnd);
            for (uint32_t w = z0; w <= z1 && z0 <= z1; ++w)
                for (uint32_t v = T.bmin[1]; v <= T.bmax[1]; ++v)
                    for (uint32_t u = T.bmin[0]; u <= T.bmax[0]; ++u) {
                        V3 uvw = {(float)u, (float)v, (float)w};
                        if (!plane_overlaps(T.v[0], T.n, offP, dP, uvw)) continue;
                        if (!tri_overlaps(T, offP, dP, uvw)) continue;
                        size_t i = VI(u, v, w);
                        counts[i] += 1u;
                        int32_t* s = &slots[4 * i];
                        if (s[0] == -1) s[0] = (int32_t)t;
                        else if (s[1] == -1) s[1] = (int32_t)t;
                        else if (s[2] == -1) s[2] = (int32_t)t;
                        else s[3] = (int32_t)t;
                        r->occ[i >> 5] |= 1u << (i & 31);
                    }
        }

        const bool textured = G.rgba8 != nullptr && G.uvs != nullptr;
        const Image im = {G.rgba8, G.width, G.height, G.addr0, G.addr1};
        std::fill(newHit.begin(), newHit.end(), 0);
        // Colour pass over the whole chunk (VZ:2506-2547).  Every thread reads
        // its own voxel before any fill write lands (fill linearization).
        for (uint32_t z = 0; z < S[2]; ++z)
            for (uint32_t y = 0; y < S[1]; ++y)
                for (uint32_t x = 0; x < S[0]; ++x) {
                    size_t i = VI(x, y, z);
                    V3 c = {comp(offP, 0) + (float)x * dP.x, comp(offP, 1) + (float)y * dP.y,
                            comp(offP, 2) + (float)z * dP.z};
                    V3 voxCenter = add(c, muls(dP, 0.5f));
                    V4 color = unpack4(colors[i]);
                    V3 normal = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
                    int32_t* s = &slots[4 * i];
                    if (s[0] < 0) {
                        colors[i] = vec4_to_color(color);  // re-quantise (identity)
                        continue;
                    }
                    newHit[i] = 1;
                    double n64[3] = {0, 0, 0}, c64[4] = {0, 0, 0, 0};
                    if (twin) {
                        for (int a = 0; a < 3; ++a) n64[a] = r->n64[3 * i + a];
                        c64[0] = (double)(colors[i] & 255u) / 255.0;
                        c64[1] = (double)((colors[i] >> 8) & 255u) / 255.0;
                        c64[2] = (double)((colors[i] >> 16) & 255u) / 255.0;
                        c64[3] = (double)(colors[i] >> 24) / 255.0;
                    }
                    if (textured) {
                        // ComputeColorsAndNormals (VK:689-918).
                        uint32_t triCount = std::min(counts[i], 4u);
                        counts[i] = 1u;
                        for (int k = 0; k < 4 && s[k] != -1; ++k) {
                            size_t ti = (size_t)s[k];
                            s[k] = -1;
                            const Tri& T = tris[ti];
                            size_t lt = ti - grpOff[g];
                            const float* uvp = G.uvs + 6 * lt;
                            V2 uv1 = {uvp[0], uvp[1]}, uv2 = {uvp[2], uvp[3]}, uv3 = {uvp[4], uvp[5]};
                            const V3& fn = T.n;
                            float ax = std::fabs(fn.x), ay = std::fabs(fn.y), az = std::fabs(fn.z);
                            V2 bc;
                            if (ax > ay && ax > az)
                                bc = bary({voxCenter.z, voxCenter.y}, {T.v[0].z, T.v[0].y},
                                          {T.v[1].z, T.v[1].y}, {T.v[2].z, T.v[2].y});
                            else if (ay > az)
                                bc = bary({voxCenter.x, voxCenter.z}, {T.v[0].x, T.v[0].z},
                                          {T.v[1].x, T.v[1].z}, {T.v[2].x, T.v[2].z});
                            else
                                bc = bary({voxCenter.x, voxCenter.y}, {T.v[0].x, T.v[0].y},
                                          {T.v[1].x, T.v[1].y}, {T.v[2].x, T.v[2].y});
                            V3 add_n;
                            if (G.vnormals != nullptr) {
                                const float* q = G.vnormals + 9 * lt;
                                add_n = interp_n(bc, {q[0], q[1], q[2]}, {q[3], q[4], q[5]},
                                                 {q[6], q[7], q[8]});
                            } else {
                                add_n = fn;
                            }
                            normal = add(normal, add_n);
                            if (twin) {
                                if (G.vnormals != nullptr) {
                                    const float* q = G.vnormals + 9 * lt;
                                    for (int a = 0; a < 3; ++a)
                                        n64[a] += (double)q[a] + ((double)q[6 + a] - q[a]) * bc.x +
                                                  ((double)q[3 + a] - q[a]) * bc.y;
                                } else {
                                    n64[0] += fn.x; n64[1] += fn.y; n64[2] += fn.z;
                                }
                            }
                            V2 tc = interp_uv(bc, uv1, uv2, uv3);
                            V4 tx = tex2d(im, tc.x, tc.y);
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
                            color.x /= (float)triCount;
                            color.y /= (float)triCount;
                            color.z /= (float)triCount;
                            color.w = gmin(color.w, 1.0f);
                            normal.x /= (float)triCount;
                            normal.y /= (float)triCount;
                            normal.z /= (float)triCount;
                            normal = normalize3(normal);
                            if (twin) {
                                double l = std::sqrt(n64[0] * n64[0] + n64[1] * n64[1] + n64[2] * n64[2]);
                                for (int a = 0; a < 3; ++a) n64[a] = n64[a] / (double)triCount;
                                l = std::sqrt(n64[0] * n64[0] + n64[1] * n64[1] + n64[2] * n64[2]);
                                for (int a = 0; a < 3; ++a) n64[a] /= l;
                            }
                        }
                    } else {
                        // ComputeColorsAsNormals (VK:920-1083).
                        uint32_t triCount = counts[i];
                        counts[i] = 1u;
                        for (int k = 0; k < 4 && s[k] != -1; ++k) {
                            size_t ti = (size_t)s[k];
                            s[k] = -1;
                            const Tri& T = tris[ti];
                            const V3& fn = T.n;
                            if (G.vnormals != nullptr) {
                                size_t lt = ti - grpOff[g];
                                const float* q = G.vnormals + 9 * lt;
                                V2 bc;
                                if (fn.x > fn.y && fn.x > fn.z)
                                    bc = bary({voxCenter.z, voxCenter.y}, {T.v[0].z, T.v[0].y},
                                              {T.v[1].z, T.v[1].y}, {T.v[2].z, T.v[2].y});
                                else if (fn.y > fn.z)
                                    bc = bary({voxCenter.x, voxCenter.z}, {T.v[0].x, T.v[0].z},
                                              {T.v[1].x, T.v[1].z}, {T.v[2].x, T.v[2].z});
                                else
                                    bc = bary({voxCenter.x, voxCenter.y}, {T.v[0].x, T.v[0].y},
                                              {T.v[1].x, T.v[1].y}, {T.v[2].x, T.v[2].y});
                                V3 ni = interp_n(bc, {q[0], q[1], q[2]}, {q[3], q[4], q[5]},
                                                 {q[6], q[7], q[8]});
                                color.x += ni.x;
                                color.y += ni.y;
                                color.z += ni.z;
                                color.w += 1.0f;
                                if (twin)
                                    for (int a = 0; a < 3; ++a)
                                        c64[a] += (double)q[a] + ((double)q[6 + a] - q[a]) * bc.x +
                                                  ((double)q[3 + a] - q[a]) * bc.y;
                            } else {
                                color.x += fn.x;
                                color.y += fn.y;
                                color.z += fn.z;
                                color.w += 1.0f;
                                if (twin) { c64[0] += fn.x; c64[1] += fn.y; c64[2] += fn.z; }
                            }
                        }
                        color.x /= (float)triCount;
                        color.y /= (float)triCount;
                        color.z /= (float)triCount;
                        color.w = gmin(color.w, 1.0f);
                        normal = normalize3(V3{color.x, color.y, color.z});
                        if (color.x < 0.0f) color.x *= -1.0f;
                        if (color.y < 0.0f) color.y *= -1.0f;
                        if (color.z < 0.0f) color.z *= -1.0f;
                        if (twin) {
                            double l = std::sqrt(c64[0] * c64[0] + c64[1] * c64[1] + c64[2] * c64[2]);
                            for (int a = 0; a < 3; ++a) n64[a] = c64[a] / l;
                        }
                    }
                    colors[i] = vec4_to_color(color);
                    normals[3 * i] = normal.x;
                    normals[3 * i + 1] = normal.y;
                    normals[3 * i + 2] = normal.z;
                    if (twin)
                        for (int a = 0; a < 3; ++a) r->n64[3 * i + a] = n64[a];
                }
        // Terrain fill (VK:886-917), textured groups only, chunk-local,
        // descending-z linearization.
        if (textured && G.is_terrain) {
            for (int64_t z = (int64_t)S[2] - 1; z >= 0; --z)
                for (uint32_t y = 0; y < S[1]; ++y)
                    for (uint32_t x = 0; x < S[0]; ++x) {
                        size_t i = VI(x, y, (size_t)z);
                        if (!newHit[i]) continue;
                        int64_t zz = z;
                        for (uint32_t gz = 1; gz < 20u; ++gz) {
                            if (zz == 0) break;
                            --zz;
                            size_t j = VI(x, y, (size_t)zz);
                            if (newHit[j]) continue;
                            colors[j] = colors[i];
                            for (int a = 0; a < 3; ++a) normals[3 * j + a] = normals[3 * i + a];
                            if (twin)
                                for (int a = 0; a < 3; ++a) r->n64[3 * j + a] = r->n64[3 * i + a];
                        }
                    }
        }
    }

    // Mip chain (computeMipMaps VZ:491-594, ComputeVoxelMipMap8 VK:1520-1663).
    r->lv.resize(L);
    r->lv[0].d[0] = S[0]; r->lv[0].d[1] = S[1]; r->lv[0].d[2] = S[2];
    r->lv[0].col.swap(colors);
    r->lv[0].nrm.swap(normals);
    for (int k = 1; k < L; ++k) {
        int src = (flags & REF_MIP_CHAIN) ? k - 1 : (k == 1 ? 0 : k - 2);
        const Level& in = r->lv[src];
        Level& out = r->lv[k];
        for (int a = 0; a < 3; ++a) out.d[a] = S[a] >> k;
        size_t n = (size_t)out.d[0] * out.d[1] * out.d[2];
        out.col.assign(n, 0u);
        out.nrm.assign(3 * n, 0.0f);
        for (uint32_t z = 0; z < out.d[2]; ++z)
            for (uint32_t y = 0; y < out.d[1]; ++y)
                for (uint32_t x = 0; x < out.d[0]; ++x) {
                    uint32_t bx = 2 * x, by = 2 * y, bz = 2 * z;
                    uint32_t x1 = std::min(bx + 1, in.d[0] - 1);
                    uint32_t y1 = std::min(by + 1, in.d[1] - 1);
                    uint32_t z1 = std::min(bz + 1, in.d[2] - 1);
                    const size_t id[8] = {in.idx(bx, by, bz), in.idx(x1, by, bz), in.idx(bx, y1, bz),
                                          in.idx(x1, y1, bz), in.idx(bx, by, z1), in.idx(x1, by, z1),
                                          in.idx(bx, y1, z1), in.idx(x1, y1, z1)};
                    V4 avg = unpack4(in.col[id[0]]);
                    V3 an = {in.nrm[3 * id[0]], in.nrm[3 * id[0] + 1], in.nrm[3 * id[0] + 2]};
                    for (int i = 1; i < 8; ++i) {
                        V4 cur = unpack4(in.col[id[i]]);
                        avg.x += cur.x; avg.y += cur.y; avg.z += cur.z; avg.w += cur.w;
                        an = add(an, V3{in.nrm[3 * id[i]], in.nrm[3 * id[i] + 1], in.nrm[3 * id[i] + 2]});
                    }
                    const float oneOverEight = 1.0f / 8.0f;
                    avg.w *= oneOverEight;
                    avg.x *= oneOverEight;
                    avg.y *= oneOverEight;
                    avg.z *= oneOverEight;
                    an = muls(an, oneOverEight);
                    float len2 = dot3(an, an);
                    if (len2 > 0.0f) {
                        float l = std::sqrt(len2);
                        an = {an.x / l, an.y / l, an.z / l};
                    } else {
                        an = {0.0f, 0.0f, 0.0f};
                    }
                    size_t o = out.idx(x, y, z);
                    out.col[o] = vec4_to_color(avg);
                    out.nrm[3 * o] = an.x;
                    out.nrm[3 * o + 1] = an.y;
                    out.nrm[3 * o + 2] = an.z;
                }
    }

    // Octree nodes, leaf to root (computeOctTreeNodes VZ:931-1290).
    r->tree.resize(L);
    uint32_t border[3] = {r->extra[0], r->extra[1], r->extra[2]};
    for (int d = L - 1; d >= 0; --d) {
        int k = L - 1 - d;
        const Level& lv = r->lv[k];
        TreeLevel& tl = r->tree[d];
        const uint32_t Rd = 1u << d;  // cube root of 8^d
        const size_t nn = (size_t)Rd * Rd * Rd;
        tl.types.assign(nn, 0);
        tl.cc.assign(nn, 0u);
        tl.bidx.assign(nn, 0xffffffffu);
        const int32_t Ck[3] = {(int32_t)lv.d[0], (int32_t)lv.d[1], (int32_t)lv.d[2]};
        const int32_t Bi[3] = {(int32_t)B[0], (int32_t)B[1], (int32_t)B[2]};
        const int32_t bo[3] = {(int32_t)border[0], (int32_t)border[1], (int32_t)border[2]};
        auto NI = [&](int32_t x, int32_t y, int32_t z) { return ((size_t)z * Rd + y) * Rd + x; };
        // Node range of the chunk (VZ:962-992).
        uint32_t nodeEnd[3];
        for (int a = 0; a < 3; ++a) {
            int32_t endVox = (Ck[a] - 1) - bo[a];
            uint32_t e = (uint32_t)(int64_t)((float)(endVox + 1) / (float)Bi[a]);
            nodeEnd[a] = e >= Rd ? Rd - 1 : e;
        }
        // ComputeOctTreeNodeConstColor (VK:1680-1772).
        for (uint32_t nz = 0; nz <= nodeEnd[2]; ++nz)
            for (uint32_t ny = 0; ny <= nodeEnd[1]; ++ny)
                for (uint32_t nx = 0; nx <= nodeEnd[0]; ++nx) {
                    int32_t sx = (int32_t)nx * Bi[0] + bo[0];
                    int32_t sy = (int32_t)ny * Bi[1] + bo[1];
                    int32_t sz = (int32_t)nz * Bi[2] + bo[2];
                    sx = std::min(std::max(sx, 0), Ck[0] - 1);
                    sy = std::min(std::max(sy, 0), Ck[1] - 1);
                    sz = std::min(std::max(sz, 0), Ck[2] - 1);
                    tl.cc[NI(nx, ny, nz)] = lv.col[lv.idx(sx, sy, sz)];
                }
        // ComputeOctTreeNodeType launch shape (VZ:1037-1120): threads per
        // axis clamped to 8, and the x axis split into two launches when the
        // voxel extent exceeds 128.
        const int32_t fullVoxDim[3] = {(int32_t)Rd * Bi[0], (int32_t)Rd * Bi[1], (int32_t)Rd * Bi[2]};
        uint32_t tpb[3];
        for (int a = 0; a < 3; ++a) {
            tpb[a] = (uint32_t)Ck[a];
            uint32_t fvo = tpb[a] + 0u - (uint32_t)bo[a];
            if (fvo > (uint32_t)fullVoxDim[a]) tpb[a] = (uint32_t)fullVoxDim[a] - (0u - (uint32_t)bo[a]) + 2u;
        }
        uint32_t th[3] = {tpb[0], tpb[1], tpb[2]}, nb[3] = {1, 1, 1};
        if (tpb[0] > 8 || tpb[1] > 8 || tpb[2] > 8) {
            for (int a = 0; a < 3; ++a) {
                th[a] = std::min(tpb[a], 8u);
                nb[a] = (uint32_t)std::ceil((float)tpb[a] / (float)th[a]);
            }
        }
        uint32_t nexec = 1;
        if (tpb[0] > 128u) {
            nb[0] >>= 1;
            if ((nb[0] & 1u) != 0u) nb[0] += 1;
            nexec = 2;
        }
        const int32_t xCover = (int32_t)(nexec * nb[0] * th[0]);
        const int32_t yCover = (int32_t)(nb[1] * th[1]);
        const int32_t zCover = (int32_t)(nb[2] * th[2]);
        auto cmp_node = [&](int32_t x, int32_t y, int32_t z, int32_t bx, int32_t by, int32_t bz) {
            // ComputeNodeType (VK:1790-1898).
            if (bx < 0 || bx >= (int32_t)Rd || by < 0 || by >= (int32_t)Rd || bz < 0 || bz >= (int32_t)Rd)
                return;
            size_t ni = NI(bx, by, bz);
            if (lv.col[lv.idx(x, y, z)] != tl.cc[ni]) tl.types[ni] = 1;
        };
        for (int32_t z = 0; z < std::min(zCover, Ck[2]); ++z)
            for (int32_t y = 0; y < std::min(yCover, Ck[1]); ++y)
                for (int32_t x = 0; x < std::min(xCover, Ck[0]); ++x) {
                    // ComputeOctTreeNodeType (VK:2216-2380).
                    int32_t f[3] = {x - bo[0], y - bo[1], z - bo[2]};
                    bool skip = false;
                    for (int a = 0; a < 3; ++a)
                        if (f[a] < -1 || f[a] > fullVoxDim[a]) skip = true;
                    if (skip) continue;
                    int32_t b[3], base[3];
                    const int32_t c3[3] = {x, y, z};
                    for (int a = 0; a < 3; ++a) {
                        b[a] = std::min((int32_t)trunc_i((float)f[a] / (float)Bi[a]), (int32_t)Rd - 1);
                        base[a] = b[a] * Bi[a] + bo[a];
                    }
                    auto lowB = [&](int a) { return c3[a] == base[a] && b[a] > 0; };
                    auto highB = [&](int a) { return c3[a] == base[a] + Bi[a] - 1 && b[a] < (int32_t)Rd - 1; };
                    auto nbr = [&](int a) { return lowB(a) ? b[a] - 1 : b[a] + 1; };
                    cmp_node(x, y, z, b[0], b[1], b[2]);
                    // HandleXBorder (VK:2082-2214) with nested Y and Z.
                    if (lowB(0) || highB(0)) {
                        int32_t bx = nbr(0);
                        cmp_node(x, y, z, bx, b[1], b[2]);
                        if (lowB(1) || highB(1)) {
                            int32_t by = nbr(1);
                            cmp_node(x, y, z, bx, by, b[2]);
                            if (lowB(2) || highB(2)) cmp_node(x, y, z, bx, by, nbr(2));
                        }
                        if (lowB(2) || highB(2)) cmp_node(x, y, z, bx, b[1], nbr(2));
                    }
                    // HandleYBorder (VK:1982-2080) with nested Z.
                    if (lowB(1) || highB(1)) {
                        int32_t by = nbr(1);
                        cmp_node(x, y, z, b[0], by, b[2]);
                        if (lowB(2) || highB(2)) cmp_node(x, y, z, b[0], by, nbr(2));
                    }
                    // HandleZBorder (VK:1900-1980).
                    if (lowB(2) || highB(2)) cmp_node(x, y, z, b[0], b[1], nbr(2));
                }
        // Brick writes in node loop order (VZ:1132-1262, BW:725-857).
        const uint32_t bd[3] = {B[0] + 2, B[1] + 2, B[2] + 2};
        const size_t bvox = (size_t)bd[0] * bd[1] * bd[2];
        for (uint32_t nz = 0; nz <= nodeEnd[2]; ++nz)
            for (uint32_t ny = 0; ny <= nodeEnd[1]; ++ny)
                for (uint32_t nx = 0; nx <= nodeEnd[0]; ++nx) {
                    int32_t off[3] = {(int32_t)nx * Bi[0] + bo[0], (int32_t)ny * Bi[1] + bo[1],
                                      (int32_t)nz * Bi[2] + bo[2]};
                    bool partial = false;
                    for (int a = 0; a < 3; ++a)
                        if (off[a] + Bi[a] >= Ck[a] || off[a] <= 0) partial = true;
                    if (partial) {
                        g_error = "brick straddles the chunk edge (single-chunk regime violated)";
                        delete r;
                        return nullptr;
                    }
                    size_t ni = NI(nx, ny, nz);
                    if (tl.types[ni] != 1) continue;
                    tl.bidx[ni] = (uint32_t)tl.nb++;
                    size_t base = tl.bcol.size();
                    tl.bcol.resize(base + bvox);
                    tl.bnrm.resize(base + bvox);
                    size_t o = base;
                    for (int32_t z = off[2] - 1; z < off[2] + Bi[2] + 1; ++z)
                        for (int32_t y = off[1] - 1; y < off[1] + Bi[1] + 1; ++y)
                            for (int32_t x = off[0] - 1; x < off[0] + Bi[0] + 1; ++x, ++o) {
                                size_t vi = lv.idx(x, y, z);
                                tl.bcol[o] = lv.col[vi];
                                float nX = ((lv.nrm[3 * vi] + 1.0f) * 0.5f);
                                float nY = ((lv.nrm[3 * vi + 1] + 1.0f) * 0.5f);
                                float nZ = ((lv.nrm[3 * vi + 2] + 1.0f) * 0.5f);
                                tl.bnrm[o] = pack4(quant_u8(nX * 255.0f), quant_u8(nY * 255.0f),
                                                   quant_u8(nZ * 255.0f), 0);
                            }
                }
        for (int a = 0; a < 3; ++a) border[a] = std::max(border[a] >> 1, 1u);
    }

    // CollapseConstantBranches (VZ:701-774): a CONST node whose whole
    // subtree is CONST becomes LEAF-CONST (2); leaves stay 0.
    for (int d = L - 2; d >= 0; --d) {
        const uint32_t Rd = 1u << d, Rc = Rd << 1;
        TreeLevel& tl = r->tree[d];
        const TreeLevel& ch = r->tree[d + 1];
        const bool childIsLeaf = (d + 1 == L - 1);
        for (uint32_t z = 0; z < Rd; ++z)
            for (uint32_t y = 0; y < Rd; ++y)
                for (uint32_t x = 0; x < Rd; ++x) {
                    bool all = true;
                    for (uint32_t cz = 2 * z; cz < 2 * z + 2; ++cz)
                        for (uint32_t cy = 2 * y; cy < 2 * y + 2; ++cy)
                            for (uint32_t cx = 2 * x; cx < 2 * x + 2; ++cx) {
                                uint8_t t = ch.types[((size_t)cz * Rc + cy) * Rc + cx];
                                bool c = childIsLeaf ? (t == 0) : (t == 2);
                                if (!c) all = false;
                            }
                    size_t i = ((size_t)z * Rd + y) * Rd + x;
                    if (tl.types[i] == 0 && all) tl.types[i] = 2;
                }
    }

    *status = anyOverlap ? 1 : -1;
    return r;
}

extern "C" int32_t ref_levels(const ref_result* r) { return r->L; }
extern "C" void ref_dims(const ref_result* r, uint32_t S[3]) {
    for (int a = 0; a < 3; ++a) S[a] = r->S[a];
}
extern "C" const uint32_t* ref_occupancy(const ref_result* r, uint64_t* nwords) {
    *nwords = r->occ.size();
    return r->occ.data();
}
extern "C" const uint8_t* ref_types(const ref_result* r, int32_t d, uint64_t* n) {
    *n = r->tree[d].types.size();
    return r->tree[d].types.data();
}
extern "C" const uint32_t* ref_const_colors(const ref_result* r, int32_t d, uint64_t* n) {
    *n = r->tree[d].cc.size();
    return r->tree[d].cc.data();
}
extern "C" const uint32_t* ref_brick_index(const ref_result* r, int32_t d, uint64_t* n) {
    *n = r->tree[d].bidx.size();
    return r->tree[d].bidx.data();
}
extern "C" uint64_t ref_num_bricks(const ref_result* r, int32_t d) { return r->tree[d].nb; }
extern "C" const uint32_t* ref_brick_colors(const ref_result* r, int32_t d) {
    return r->tree[d].bcol.data();
}
extern "C" const uint32_t* ref_brick_normals(const ref_result* r, int32_t d) {
    return r->tree[d].bnrm.data();
}
extern "C" const uint32_t* ref_level_colors(const ref_result* r, int32_t k, uint32_t dims[3]) {
    for (int a = 0; a < 3; ++a) dims[a] = r->lv[k].d[a];
    return r->lv[k].col.data();
}
extern "C" const float* ref_level_normals(const ref_result* r, int32_t k, uint32_t dims[3]) {
    for (int a = 0; a < 3; ++a) dims[a] = r->lv[k].d[a];
    return r->lv[k].nrm.data();
}
extern "C" void ref_free(ref_result* r) { delete r; }

extern "C" double ref_normal_error_fp64(const ref_result* r, uint64_t* nvox) {
    *nvox = 0;
    if (!r->has64) return -1.0;
    const Level& lv = r->lv[0];
    double worst = 0.0;
    size_t n = lv.col.size();
    for (size_t i = 0; i < n; ++i) {
        double a[3] = {lv.nrm[3 * i], lv.nrm[3 * i + 1], lv.nrm[3 * i + 2]};
        double b[3] = {r->n64[3 * i], r->n64[3 * i + 1], r->n64[3 * i + 2]};
        double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        if (!(la > 0.5 && la < 1.5) || !(lb > 0.5 && lb < 1.5)) continue;
        double c = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (la * lb);
        c = std::min(1.0, std::max(-1.0, c));
        double ang = std::acos(c);
        // acos loses precision near 1; use the cross-product magnitude.
        double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
        double s = std::sqrt(cx * cx + cy * cy + cz * cz) / (la * lb);
        ang = std::atan2(s, c);
        worst = std::max(worst, ang);
        ++*nvox;
    }
    return worst;
}
