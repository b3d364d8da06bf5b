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
#include <array>
#include <deque>
#include <map>
#include <set>
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
    uint32_t C[3], nc[3];  // chunk edge and chunks per axis
    int L = 0;
    std::vector<uint32_t> occ;
    std::vector<Level> lv;        // mip level k (0 = full res)
    std::vector<TreeLevel> tree;  // tree depth d (0 = root)
    std::vector<double> n64;      // fp64 twin normals at level 0 (optional)
    bool has64 = false;
};

extern "C" const char* ref_error(void) { return g_error.c_str(); }

namespace {

// A brick that straddles chunks while it is assembled
// (VoxelBrickWriter::BrickData, BW:126-205 and 862-1064): (B+2)^3 colours and
// float normals, and the number of voxels stored so far.  It is complete when
// that number reaches (B+2)^3 (VoxelBrickWriter.h 78-81).
struct PartialBrick {
    std::vector<uint32_t> col;
    std::vector<float> nrm;  // 3 per voxel
    uint64_t stored = 0;
};

// Node / chunk key ordered z, then y, then x (BrickID, VoxelBrickWriter.h
// 20-31; VoxelChunkID, VZ:2159-2173).
struct Key3 {
    uint32_t x, y, z;
    bool operator<(const Key3& o) const {
        return z < o.z || (z == o.z && y < o.y) || (z == o.z && y == o.y && x < o.x);
    }
};

// Append one brick to the depth's stream with the next index (writeBrick,
// BW:820-857: index = _brickMap.size()); colours as stored, normals written
// as (n + 1) / 2 * 255 truncated (WriteUByteBinaryBrick, BW:509-588).
template <class ColAt, class NrmAt>
void write_brick(TreeLevel& tl, size_t node, size_t bvox, ColAt colAt, NrmAt nrmAt) {
    tl.bidx[node] = (uint32_t)tl.nb++;
    size_t base = tl.bcol.size();
    tl.bcol.resize(base + bvox);
    tl.bnrm.resize(base + bvox);
    for (size_t o = 0; o < bvox; ++o) {
        tl.bcol[base + o] = colAt(o);
        const float* n = nrmAt(o);
        float nX = ((n[0] + 1.0f) * 0.5f);
        float nY = ((n[1] + 1.0f) * 0.5f);
        float nZ = ((n[2] + 1.0f) * 0.5f);
        tl.bnrm[base + o] = pack4(quant_u8(nX * 255.0f), quant_u8(nY * 255.0f),
                                  quant_u8(nZ * 255.0f), 0);
    }
}

}  // namespace

extern "C" ref_result* ref_run_chunked(const uint32_t N[3], const uint32_t B[3],
                                       const uint32_t chunk[3], const float pin[3],
                                       const float delta[3], const ref_group* groups,
                                       int32_t ngroups, int32_t flags, int32_t* status) {
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
    // Chunk edge per axis (voxChunkDim, VH:124, VZ:2184).  A chunk at least S
    // wide holds the whole voxel space, and the run is then the same as with
    // a chunk of exactly S (every brick lies inside it and the typing launch
    // covers the same voxels), so S is used.  A smaller chunk must be a
    // multiple of 2^(L-1), so that the chunk origin of every level (x >> k,
    // VZ:1273-1276) is a whole number of that level's chunk edges.
    uint32_t Ce[3], nc[3];
    for (int a = 0; a < 3; ++a) {
        if (!chunk || chunk[a] >= S[a]) {
            Ce[a] = S[a];
        } else {
            if (chunk[a] == 0 || chunk[a] % (1u << (L - 1)) != 0) {
                g_error = "chunk must be a multiple of 2^(levels-1) or at least N + 4*((N/B)>>1)";
                delete r;
                return nullptr;
            }
            Ce[a] = chunk[a];
        }
        // chunk origins are the multiples of Ce below N (VZ:2300-2302)
        nc[a] = (N[a] + Ce[a] - 1) / Ce[a];
        r->C[a] = Ce[a];
        r->nc[a] = nc[a];
    }
    const V3 dP = {delta[0], delta[1], delta[2]};
    const V3 P = {pin[0], pin[1], pin[2]};
    // setVoxelizationParams (VZ:155-161): offsetP = p - extra*delta.
    const V3 offP = sub(P, mulv(V3{(float)r->extra[0], (float)r->extra[1], (float)r->extra[2]}, dP));
    const bool twin = (flags & REF_FP64_TWIN) != 0;

    // Chunk visit order: breadth first from the origin, each chunk queueing
    // its +x/+y/+z neighbours in x, y, z loop order (VZ:2285-2309).
    std::vector<Key3> order;
    {
        std::set<Key3> queued;
        std::deque<Key3> q;
        q.push_back({0, 0, 0});
        queued.insert(q.front());
        while (!q.empty()) {
            Key3 c = q.front();
            q.pop_front();
            order.push_back(c);
            for (uint32_t nx = c.x; nx <= c.x + Ce[0]; nx += Ce[0])
                for (uint32_t ny = c.y; ny <= c.y + Ce[1]; ny += Ce[1])
                    for (uint32_t nz = c.z; nz <= c.z + Ce[2]; nz += Ce[2]) {
                        Key3 n = {nx, ny, nz};
                        if (nx < N[0] && ny < N[1] && nz < N[2] && queued.insert(n).second)
                            q.push_back(n);
                    }
        }
    }

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
    // Group voxel bounds (addTriangleGroup, VH:137-144): clamp to [0, S], cast.
    std::vector<uint32_t> gmn(3 * ngroups), gmx(3 * ngroups);
    for (int g = 0; g < ngroups; ++g)
        for (int a = 0; a < 3; ++a) {
            float lo = gclamp((groups[g].bbox_min[a] - comp(offP, a)) / comp(dP, a), 0.0f, (float)S[a]);
            float hi = gclamp((groups[g].bbox_max[a] - comp(offP, a)) / comp(dP, a), 0.0f, (float)S[a]);
            gmn[3 * g + a] = (uint32_t)trunc_i(lo);
            gmx[3 * g + a] = (uint32_t)trunc_i(hi);
        }

    // Outputs.  Level k is laid out chunk by chunk: chunk (i, j, l) holds
    // voxels [i, i+1) * (Ce >> k) x ... of it, so that with one chunk it is
    // the voxel space itself.
    r->occ.assign(((size_t)S[0] * S[1] * S[2] + 31) / 32, 0u);
    r->lv.resize(L);
    for (int k = 0; k < L; ++k) {
        for (int a = 0; a < 3; ++a) r->lv[k].d[a] = nc[a] * (Ce[a] >> k);
        size_t n = (size_t)r->lv[k].d[0] * r->lv[k].d[1] * r->lv[k].d[2];
        r->lv[k].col.assign(n, 0u);
        r->lv[k].nrm.assign(3 * n, 0.0f);
    }
    if (twin) { r->n64.assign(r->lv[0].nrm.size(), 0.0); r->has64 = true; }
    // Node buffers of each depth on the device (initOctTreeDeviceBuffers,
    // VZ:895-929: zeroed once per generate) and their host copies (VZ:2208-2213).
    std::vector<std::vector<uint32_t>> devT(L), devC(L), hostT(L), hostC(L);
    std::vector<std::map<Key3, PartialBrick>> parts(L);
    r->tree.resize(L);
    for (int d = 0; d < L; ++d) {
        const size_t nn = (size_t)1 << (3 * d);
        devT[d].assign(nn, 0u);
        devC[d].assign(nn, 0u);
        hostT[d].assign(nn, 0u);
        hostC[d].assign(nn, 0u);
        r->tree[d].bidx.assign(nn, 0xffffffffu);
    }
    // Brick border of each level (VZ:1270-1272): extra at level 0, then
    // halved, at least 1.
    std::vector<std::array<uint32_t, 3>> bord(L);
    for (int a = 0; a < 3; ++a) bord[0][a] = r->extra[a];
    for (int k = 1; k < L; ++k)
        for (int a = 0; a < 3; ++a) bord[k][a] = std::max(bord[k - 1][a] >> 1, 1u);
    const uint32_t bd[3] = {B[0] + 2, B[1] + 2, B[2] + 2};
    const size_t bvox = (size_t)bd[0] * bd[1] * bd[2];

    // Per-chunk state (allocateDeviceVoxelizationMemory VZ:1349-1411: slots
    // start at -1 and every colour pass resets the ones it read).
    const size_t cvox = (size_t)Ce[0] * Ce[1] * Ce[2];
    std::vector<uint32_t> counts(cvox);
    std::vector<int32_t> slots(4 * cvox, -1);
    std::vector<uint32_t> colors(cvox);
    std::vector<float> normals(3 * cvox);
    std::vector<uint8_t> newHit(cvox);
    std::vector<double> c64;
    if (twin) c64.assign(3 * cvox, 0.0);
    auto VI = [&](size_t x, size_t y, size_t z) { return (z * Ce[1] + y) * Ce[0] + x; };
    std::vector<Level> clv(L);

    bool anyOverlap = false;
    for (const Key3& ch : order) {
        const uint32_t o0[3] = {ch.x, ch.y, ch.z};
        const uint32_t ci[3] = {ch.x / Ce[0], ch.y / Ce[1], ch.z / Ce[2]};
        // memsets at the start of each chunk (VZ:2325-2332)
        // (the previous chunk's level 0 was moved into clv[0])
        std::fill(counts.begin(), counts.end(), 0u);
        colors.assign(cvox, 0u);
        normals.assign(3 * cvox, 0.0f);
        if (twin) std::fill(c64.begin(), c64.end(), 0.0);
        // voxOffsetOrigin (VZ:2317-2319), in float on the host
        const V3 org = {offP.x + (float)o0[0] * dP.x, offP.y + (float)o0[1] * dP.y,
                        offP.z + (float)o0[2] * dP.z};
        for (int g = 0; g < ngroups; ++g) {
            const ref_group& G = groups[g];
            // Chunk overlap test (VZ:2341-2353).
            bool overlaps = true;
            for (int a = 0; a < 3; ++a)
                if (gmn[3 * g + a] > o0[a] + Ce[a] - 1 || gmx[3 * g + a] < o0[a]) overlaps = false;
            if (!overlaps) continue;
            anyOverlap = true;
            const uint32_t zStart = std::max(gmn[3 * g + 2], o0[2]);
            const uint32_t zEnd = std::min(gmx[3 * g + 2], o0[2] + Ce[2] - 1);

            // ComputeVoxelization (VK:349-530), one triangle at a time in
            // index order (the cpuTest loop shape, VK:2426-2505), over the
            // triangle's bounds clipped to the chunk.
            for (size_t t = grpOff[g]; t < grpOff[g + 1]; ++t) {
                const Tri& T = tris[t];
                uint32_t lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::max(T.bmin[a], o0[a]);
                    hi[a] = std::min(T.bmax[a], o0[a] + Ce[a] - 1);
                }
                lo[2] = std::max(lo[2], zStart);
                hi[2] = std::min(hi[2], zEnd);
                if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) continue;
                for (uint32_t w = lo[2]; w <= hi[2]; ++w)
                    for (uint32_t v = lo[1]; v <= hi[1]; ++v)
                        for (uint32_t u = lo[0]; u <= hi[0]; ++u) {
                            V3 uvw = {(float)u, (float)v, (float)w};
                            if (!plane_overlaps(T.v[0], T.n, offP, dP, uvw)) continue;
                            if (!tri_overlaps(T, offP, dP, uvw)) continue;
                            size_t i = VI(u - o0[0], v - o0[1], w - o0[2]);
                            counts[i] += 1u;
                            int32_t* s = &slots[4 * i];
                            if (s[0] == -1) s[0] = (int32_t)t;
                            else if (s[1] == -1) s[1] = (int32_t)t;
                            else if (s[2] == -1) s[2] = (int32_t)t;
                            else s[3] = (int32_t)t;
                            size_t gi = ((size_t)w * S[1] + v) * S[0] + u;
                            r->occ[gi >> 5] |= 1u << (gi & 31);
                        }
            }

            const bool textured = G.rgba8 != nullptr && G.uvs != nullptr;
            const Image im = {G.rgba8, G.width, G.height, G.addr0, G.addr1};
            std::fill(newHit.begin(), newHit.end(), 0);
            // Colour pass over the whole chunk (VZ:2506-2547).  Every thread
            // reads its own voxel before any fill write lands (fill
            // linearization).
            for (uint32_t z = 0; z < Ce[2]; ++z)
                for (uint32_t y = 0; y < Ce[1]; ++y)
                    for (uint32_t x = 0; x < Ce[0]; ++x) {
                        size_t i = VI(x, y, z);
                        // voxel centre (VK:715-719): origin + index * size + size / 2
                        V3 c = {org.x + (float)x * dP.x, org.y + (float)y * dP.y, org.z + (float)z * dP.z};
                        V3 voxCenter = add(c, muls(dP, 0.5f));
                        V4 color = unpack4(colors[i]);
                        V3 normal = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
                        int32_t* s = &slots[4 * i];
                        if (s[0] < 0) {
                            colors[i] = vec4_to_color(color);  // re-quantise (identity)
                            continue;
                        }
                        newHit[i] = 1;
                        double n64[3] = {0, 0, 0}, cc64[4] = {0, 0, 0, 0};
                        if (twin)
                            for (int a = 0; a < 3; ++a) n64[a] = c64[3 * i + a];
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
                                    for (int a = 0; a < 3; ++a) n64[a] = n64[a] / (double)triCount;
                                    double l = std::sqrt(n64[0] * n64[0] + n64[1] * n64[1] + n64[2] * n64[2]);
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
                                            cc64[a] += (double)q[a] + ((double)q[6 + a] - q[a]) * bc.x +
                                                       ((double)q[3 + a] - q[a]) * bc.y;
                                } else {
                                    color.x += fn.x;
                                    color.y += fn.y;
                                    color.z += fn.z;
                                    color.w += 1.0f;
                                    if (twin) { cc64[0] += fn.x; cc64[1] += fn.y; cc64[2] += fn.z; }
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
                                double l = std::sqrt(cc64[0] * cc64[0] + cc64[1] * cc64[1] + cc64[2] * cc64[2]);
                                for (int a = 0; a < 3; ++a) n64[a] = cc64[a] / l;
                            }
                        }
                        colors[i] = vec4_to_color(color);
                        normals[3 * i] = normal.x;
                        normals[3 * i + 1] = normal.y;
                        normals[3 * i + 2] = normal.z;
                        if (twin)
                            for (int a = 0; a < 3; ++a) c64[3 * i + a] = n64[a];
                    }
            // Terrain fill (VK:886-917), textured groups only, chunk-local
            // (it stops at the chunk's z = 0), descending-z linearization.
            if (textured && G.is_terrain) {
                for (int64_t z = (int64_t)Ce[2] - 1; z >= 0; --z)
                    for (uint32_t y = 0; y < Ce[1]; ++y)
                        for (uint32_t x = 0; x < Ce[0]; ++x) {
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
                                    for (int a = 0; a < 3; ++a) c64[3 * j + a] = c64[3 * i + a];
                            }
                        }
            }
        }

        // Mip chain of the chunk (computeMipMaps VZ:491-594,
        // ComputeVoxelMipMap8 VK:1520-1663).
        clv[0].d[0] = Ce[0]; clv[0].d[1] = Ce[1]; clv[0].d[2] = Ce[2];
        clv[0].col.swap(colors);
        clv[0].nrm.swap(normals);
        for (int k = 1; k < L; ++k) {
            int src = (flags & REF_MIP_CHAIN) ? k - 1 : (k == 1 ? 0 : k - 2);
            const Level& in = clv[src];
            Level& out = clv[k];
            for (int a = 0; a < 3; ++a) out.d[a] = Ce[a] >> k;
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
        // The chunk's levels into the chunk-by-chunk output layout.
        for (int k = 0; k < L; ++k) {
            const Level& in = clv[k];
            Level& out = r->lv[k];
            for (uint32_t z = 0; z < in.d[2]; ++z)
                for (uint32_t y = 0; y < in.d[1]; ++y) {
                    size_t si = in.idx(0, y, z);
                    size_t di = out.idx(ci[0] * in.d[0], ci[1] * in.d[1] + y, ci[2] * in.d[2] + z);
                    std::copy(in.col.begin() + si, in.col.begin() + si + in.d[0], out.col.begin() + di);
                    std::copy(in.nrm.begin() + 3 * si, in.nrm.begin() + 3 * (si + in.d[0]),
                              out.nrm.begin() + 3 * di);
                    if (k == 0 && twin)
                        std::copy(c64.begin() + 3 * si, c64.begin() + 3 * (si + in.d[0]),
                                  r->n64.begin() + 3 * di);
                }
        }

        // Octree nodes of the chunk, leaf to root (computeOctTreeNodes
        // VZ:931-1290): chunk origin and edge halve per level.
        for (int d = L - 1; d >= 0; --d) {
            const int k = L - 1 - d;
            const Level& lv = clv[k];
            const uint32_t Rd = 1u << d;  // cube root of 8^d
            auto NI = [&](int32_t x, int32_t y, int32_t z) { return ((size_t)z * Rd + y) * Rd + x; };
            const int32_t Ck[3] = {(int32_t)lv.d[0], (int32_t)lv.d[1], (int32_t)lv.d[2]};
            const int32_t ok[3] = {(int32_t)(o0[0] >> k), (int32_t)(o0[1] >> k), (int32_t)(o0[2] >> k)};
            const int32_t Bi[3] = {(int32_t)B[0], (int32_t)B[1], (int32_t)B[2]};
            const int32_t bo[3] = {(int32_t)bord[k][0], (int32_t)bord[k][1], (int32_t)bord[k][2]};
            std::vector<uint32_t>& dT = devT[d];
            std::vector<uint32_t>& dC = devC[d];
            // Node range of the chunk (VZ:956-992).
            uint32_t ns[3], ne[3];
            for (int a = 0; a < 3; ++a) {
                int32_t startVox = ok[a] - bo[a];
                ns[a] = 0;
                if (startVox >= 0)
                    ns[a] = (uint32_t)(int64_t)((float)std::max(startVox - 1, 0) / (float)Bi[a]);
                int32_t endVox = (ok[a] + Ck[a] - 1) - bo[a];
                uint32_t e = (uint32_t)(int64_t)((float)(endVox + 1) / (float)Bi[a]);
                ne[a] = e >= Rd ? Rd - 1 : e;
            }
            // ComputeOctTreeNodeConstColor (VK:1680-1772): the brick's first
            // interior voxel, clamped into the chunk.
            for (uint32_t nz = ns[2]; nz <= ne[2]; ++nz)
                for (uint32_t ny = ns[1]; ny <= ne[1]; ++ny)
                    for (uint32_t nx = ns[0]; nx <= ne[0]; ++nx) {
                        const uint32_t n3[3] = {nx, ny, nz};
                        int32_t sp[3];
                        for (int a = 0; a < 3; ++a) {
                            sp[a] = (int32_t)n3[a] * Bi[a] - ok[a] + bo[a];
                            sp[a] = std::min(std::max(sp[a], 0), Ck[a] - 1);
                        }
                        dC[NI(nx, ny, nz)] = lv.col[lv.idx(sp[0], sp[1], sp[2])];
                    }
            // ComputeOctTreeNodeType launch shape (VZ:1037-1120): threads per
            // axis clamped to 8, and the x axis split into two launches when
            // the voxel extent exceeds 128.  Unsigned arithmetic as there.
            const int32_t fullVoxDim[3] = {(int32_t)Rd * Bi[0], (int32_t)Rd * Bi[1], (int32_t)Rd * Bi[2]};
            uint32_t tpb[3];
            for (int a = 0; a < 3; ++a) {
                tpb[a] = (uint32_t)Ck[a];
                uint32_t fvo = tpb[a] + (uint32_t)ok[a] - (uint32_t)bo[a];
                if (fvo > (uint32_t)fullVoxDim[a])
                    tpb[a] = (uint32_t)fullVoxDim[a] - ((uint32_t)ok[a] - (uint32_t)bo[a]) + 2u;
            }
            uint32_t th[3] = {tpb[0], tpb[1], tpb[2]}, nbk[3] = {1, 1, 1};
            if (tpb[0] > 8 || tpb[1] > 8 || tpb[2] > 8) {
                for (int a = 0; a < 3; ++a) {
                    th[a] = std::min(tpb[a], 8u);
                    nbk[a] = (uint32_t)std::ceil((float)tpb[a] / (float)th[a]);
                }
            }
            uint32_t nexec = 1;
            if (tpb[0] > 128u) {
                nbk[0] >>= 1;
                if ((nbk[0] & 1u) != 0u) nbk[0] += 1;
                nexec = 2;
            }
            const int64_t xCover = (int64_t)nexec * nbk[0] * th[0];
            const int64_t yCover = (int64_t)nbk[1] * th[1];
            const int64_t zCover = (int64_t)nbk[2] * th[2];
            auto cmp_node = [&](int32_t x, int32_t y, int32_t z, int32_t bx, int32_t by, int32_t bz) {
                // ComputeNodeType (VK:1790-1898).
                if (bx < 0 || bx >= (int32_t)Rd || by < 0 || by >= (int32_t)Rd || bz < 0 || bz >= (int32_t)Rd)
                    return;
                size_t ni = NI(bx, by, bz);
                if (lv.col[lv.idx(x, y, z)] != dC[ni]) dT[ni] = 1u;
            };
            for (int32_t z = 0; z < (int32_t)std::min<int64_t>(zCover, Ck[2]); ++z)
                for (int32_t y = 0; y < (int32_t)std::min<int64_t>(yCover, Ck[1]); ++y)
                    for (int32_t x = 0; x < (int32_t)std::min<int64_t>(xCover, Ck[0]); ++x) {
                        // ComputeOctTreeNodeType (VK:2216-2380).
                        int32_t f[3] = {x + ok[0] - bo[0], y + ok[1] - bo[1], z + ok[2] - bo[2]};
                        bool skip = false;
                        for (int a = 0; a < 3; ++a)
                            if (f[a] < -1 || f[a] > fullVoxDim[a]) skip = true;
                        if (skip) continue;
                        int32_t b[3], base[3];
                        const int32_t c3[3] = {x, y, z};
                        for (int a = 0; a < 3; ++a) {
                            b[a] = std::min((int32_t)trunc_i((float)f[a] / (float)Bi[a]), (int32_t)Rd - 1);
                            base[a] = b[a] * Bi[a] - ok[a] + bo[a];
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
            // Row copies of the node range to the host arrays (VZ:1150-1172).
            for (uint32_t nz = ns[2]; nz <= ne[2]; ++nz)
                for (uint32_t ny = ns[1]; ny <= ne[1]; ++ny)
                    for (uint32_t nx = ns[0]; nx <= ne[0]; ++nx) {
                        size_t ni = NI(nx, ny, nz);
                        hostT[d][ni] = dT[ni];
                        hostC[d][ni] = dC[ni];
                    }
            // Bricks of the node range (VZ:1132-1262): a brick inside the
            // chunk with its border is written at once when NON-CONST; one
            // that reaches the chunk's edge is stored piecewise
            // (storePartialBrick / CopyPartialBrick, BW:1026-1064, 126-205).
            TreeLevel& tl = r->tree[d];
            for (uint32_t nz = ns[2]; nz <= ne[2]; ++nz)
                for (uint32_t ny = ns[1]; ny <= ne[1]; ++ny)
                    for (uint32_t nx = ns[0]; nx <= ne[0]; ++nx) {
                        const uint32_t n3[3] = {nx, ny, nz};
                        int32_t off[3];
                        bool partial = false;
                        for (int a = 0; a < 3; ++a) {
                            off[a] = (int32_t)n3[a] * Bi[a] - ok[a] + bo[a];
                            if (off[a] + Bi[a] >= Ck[a] || off[a] <= 0) partial = true;
                        }
                        size_t ni = NI(nx, ny, nz);
                        if (!partial) {
                            if (hostT[d][ni] != 1u) continue;
                            write_brick(
                                tl, ni, bvox,
                                [&](size_t o) {
                                    size_t x = o % bd[0], y = (o / bd[0]) % bd[1], z = o / ((size_t)bd[0] * bd[1]);
                                    return lv.col[lv.idx(off[0] - 1 + x, off[1] - 1 + y, off[2] - 1 + z)];
                                },
                                [&](size_t o) {
                                    size_t x = o % bd[0], y = (o / bd[0]) % bd[1], z = o / ((size_t)bd[0] * bd[1]);
                                    return &lv.nrm[3 * lv.idx(off[0] - 1 + x, off[1] - 1 + y, off[2] - 1 + z)];
                                });
                            continue;
                        }
                        // the piece of this chunk: position in the brick (with
                        // border), start in the chunk, interior size
                        int64_t bp[3], mm[3], sz[3], end[3];
                        for (int a = 0; a < 3; ++a) {
                            if (off[a] < 0) {
                                bp[a] = 1 - off[a];
                                mm[a] = 0;
                                sz[a] = Bi[a] - bp[a] + 1;
                            } else {
                                bp[a] = 1;
                                mm[a] = off[a];
                                sz[a] = std::min<int64_t>(Bi[a], (int64_t)Ck[a] - off[a]);
                            }
                            // border voxels on the sides that are not the
                            // chunk's edge (CopyPartialBrick, BW:136-173)
                            end[a] = mm[a] + sz[a];
                            if (mm[a] != 0) { mm[a] -= 1; bp[a] -= 1; }
                            if (end[a] < Ck[a]) end[a] += 1;
                            else if (end[a] > Ck[a]) end[a] = Ck[a];
                        }
                        PartialBrick& pb = parts[d][Key3{nx, ny, nz}];
                        if (pb.col.empty()) {
                            pb.col.assign(bvox, 0u);
                            pb.nrm.assign(3 * bvox, 0.0f);
                        }
                        for (int64_t z = mm[2], bz = bp[2]; z < end[2]; ++z, ++bz)
                            for (int64_t y = mm[1], by = bp[1]; y < end[1]; ++y, ++by)
                                for (int64_t x = mm[0], bx = bp[0]; x < end[0]; ++x, ++bx) {
                                    size_t src = lv.idx(x, y, z);
                                    size_t dst = ((size_t)bz * bd[1] + by) * bd[0] + bx;
                                    pb.col[dst] = lv.col[src];
                                    for (int a = 0; a < 3; ++a) pb.nrm[3 * dst + a] = lv.nrm[3 * src + a];
                                }
                        pb.stored += (uint64_t)((end[0] - mm[0]) * (end[1] - mm[1]) * (end[2] - mm[2]));
                    }
        }
        // Completed stored bricks of every depth, root first, each depth in
        // (z, y, x) order; written when NON-CONST (WriteCompletedStoredBricks,
        // VZ:1700-1720, BW:1131-1210).
        for (int d = 0; d < L; ++d) {
            const uint32_t Rd = 1u << d;
            for (auto it = parts[d].begin(); it != parts[d].end();) {
                if (it->second.stored != bvox) { ++it; continue; }
                size_t ni = ((size_t)it->first.z * Rd + it->first.y) * Rd + it->first.x;
                if (hostT[d][ni] == 1u) {
                    const PartialBrick& pb = it->second;
                    write_brick(r->tree[d], ni, bvox, [&](size_t o) { return pb.col[o]; },
                                [&](size_t o) { return &pb.nrm[3 * o]; });
                }
                it = parts[d].erase(it);
            }
        }
    }

    for (int d = 0; d < L; ++d) {
        TreeLevel& tl = r->tree[d];
        tl.types.resize(hostT[d].size());
        for (size_t i = 0; i < hostT[d].size(); ++i) tl.types[i] = (uint8_t)hostT[d][i];
        tl.cc = hostC[d];
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

extern "C" ref_result* ref_run(const uint32_t N[3], const uint32_t B[3], const float pin[3],
                               const float delta[3], const ref_group* groups, int32_t ngroups,
                               int32_t flags, int32_t* status) {
    return ref_run_chunked(N, B, nullptr, pin, delta, groups, ngroups, flags, status);
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
