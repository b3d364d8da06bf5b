// vox_pool.hip -- the GigaVoxels node pool and brick pool built on the device
// (C ABI in include/voxmi_pool.h).
//
// The reference loads tree.gvx depth-first on the host and appends a block of
// 8 nodes for every node that has children (GigaVoxelsReader.cpp 277-372,
// GigaVoxelsOctTreeNodePool.cpp 49-74), so the block of a node's children is
// 1 + its rank among the nodes with children in depth-first preorder.  Here
// that rank comes from two passes over the dense per-depth node arrays: a
// bottom-up count of the nodes with children in every subtree, then a
// top-down prefix over the 8 children of each node.  The texture size uses
// the same double-precision cube root (GigaVoxelsOctTreeNodePool.cpp
// 101-116), so texel positions and child pointers equal the reference's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/voxmi_pool.h"
#include "vox_internal.h"

using namespace voxmi;

namespace {

constexpr uint32_t kNoTexel = 0xffffffffu;
constexpr uint32_t kConstBit = 0x40000000u;
constexpr uint32_t kResidentBit = 0x80000000u;
constexpr int kPoolBlock = 256;

__device__ __forceinline__ uint32_t pack10(uint32_t x, uint32_t y, uint32_t z) {
    return (x << 20) | (y << 10) | z;
}

// Node n of a depth with 2^sh nodes per axis, x fastest.
__device__ __forceinline__ void node_xyz(uint64_t n, uint32_t sh, uint32_t& x, uint32_t& y,
                                         uint32_t& z) {
    const uint64_t m = (1ull << sh) - 1ull;
    x = (uint32_t)(n & m);
    y = (uint32_t)((n >> sh) & m);
    z = (uint32_t)(n >> (2 * sh));
}
__device__ __forceinline__ uint64_t node_id(uint32_t x, uint32_t y, uint32_t z, uint32_t sh) {
    return ((uint64_t)z << (2 * sh)) | ((uint64_t)y << sh) | x;
}
// Child c (x fastest within the 2x2x2 block, the .gvx z, y, x loop order).
__device__ __forceinline__ uint64_t child_id(uint32_t x, uint32_t y, uint32_t z, uint32_t sh,
                                             int c) {
    return node_id(2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + (c >> 2), sh + 1);
}
// Texel origin of block b: 2 * (b % bx, (b / bx) % by, b / (bx * by))
// (GigaVoxelsOctTreeNodePool.cpp 158-166).
__device__ __forceinline__ void block_origin(uint32_t b, uint32_t bx, uint32_t by, uint32_t& ox,
                                             uint32_t& oy, uint32_t& oz) {
    const uint32_t bz = b / (bx * by);
    const uint32_t r = b - bz * bx * by;
    const uint32_t byy = r / bx;
    ox = 2 * (r - byy * bx);
    oy = 2 * byy;
    oz = 2 * bz;
}

// A NON-CONST node without a brick (several chunks: its brick reaches a
// chunk that is not processed) reads brick 0, as the reader does with the
// tree file's Brick="0" (GigaVoxelsReader.cpp 316-321); a depth without any
// brick makes such a node constant instead (its stored colour).
__global__ void k_missing_bricks(uint64_t n, uint64_t nb, uint8_t* __restrict__ types,
                                 uint32_t* __restrict__ bidx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || types[i] != 1u || bidx[i] != 0xffffffffu) return;
    if (nb > 0) bidx[i] = 0u;
    else types[i] = 0u;
}

// state bit 0: node is in the tree; bit 1: node has a child block (in the tree,
// above the last depth, not LEAF-CONST: LoadOctTreeNode 326-327).
__global__ void k_tree_state(uint64_t n, uint32_t sh, int last, const uint8_t* __restrict__ types,
                             const uint8_t* __restrict__ parentState, uint8_t* __restrict__ state) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        node_xyz(i, sh, x, y, z);
        const bool in = parentState == nullptr ||
                        (parentState[node_id(x >> 1, y >> 1, z >> 1, sh - 1)] & 2u) != 0;
        const bool kids = in && !last && types[i] != 2;
        state[i] = (uint8_t)((in ? 1u : 0u) | (kids ? 2u : 0u));
    }
}

// Nodes with children in the subtree of each node (bottom-up).
__global__ void k_tree_subtree(uint64_t n, uint32_t sh, const uint8_t* __restrict__ state,
                               const uint32_t* __restrict__ childSub, uint32_t* __restrict__ sub) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t s = 0;
        if (state[i] & 2u) {
            uint32_t x, y, z;
            node_xyz(i, sh, x, y, z);
            s = 1;
            for (int c = 0; c < 8; ++c) s += childSub[child_id(x, y, z, sh, c)];
        }
        sub[i] = s;
    }
}

// Depth-first preorder rank among nodes with children (top-down): the first
// child follows its parent, each later child follows its elder siblings'
// subtrees.
__global__ void k_tree_rank(uint64_t n, uint32_t sh, const uint8_t* __restrict__ state,
                            const uint32_t* __restrict__ pre, const uint32_t* __restrict__ childSub,
                            uint32_t* __restrict__ childPre) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (!(state[i] & 2u)) continue;
        uint32_t x, y, z;
        node_xyz(i, sh, x, y, z);
        uint32_t r = pre[i] + 1;
        for (int c = 0; c < 8; ++c) {
            const uint64_t cid = child_id(x, y, z, sh, c);
            childPre[cid] = r;
            r += childSub[cid];
        }
    }
}

// Per-channel map of a constant colour into word 1.
struct ConstLut {
    uint8_t v[256];
};

// Node texels: position inside the parent's child block, word 0 / word 1
// (Node::setNodeTypeFlag, setChildNodesPtr; create3DTexture 167-188).
__global__ void k_tree_write(uint64_t n, uint32_t sh, const uint8_t* __restrict__ state,
                             const uint8_t* __restrict__ types, const uint32_t* __restrict__ cc,
                             const uint32_t* __restrict__ pre,
                             const uint32_t* __restrict__ parentPre, uint32_t bx, uint32_t by,
                             uint32_t dx, uint32_t dy, ConstLut lut, int storeConst,
                             uint32_t* __restrict__ tex, uint32_t* __restrict__ texel) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t st = state[i];
        if (!(st & 1u)) {
            texel[i] = kNoTexel;
            continue;
        }
        uint32_t x, y, z;
        node_xyz(i, sh, x, y, z);
        uint32_t tx = 0, ty = 0, tz = 0;  // the root: node 0 of block 0
        if (parentPre != nullptr) {
            const uint32_t b = 1u + parentPre[node_id(x >> 1, y >> 1, z >> 1, sh - 1)];
            block_origin(b, bx, by, tx, ty, tz);
            tx += x & 1u;
            ty += y & 1u;
            tz += z & 1u;
        }
        const bool isConst = types[i] != 1;
        uint32_t w0 = isConst ? kConstBit : 0u;
        if (st & 2u) {
            uint32_t ox, oy, oz;
            block_origin(1u + pre[i], bx, by, ox, oy, oz);
            w0 |= pack10(ox, oy, oz);
        }
        uint32_t w1 = 0;
        if (isConst && storeConst) {
            const uint32_t c = cc[i];
            w1 = (uint32_t)lut.v[c & 255u] | ((uint32_t)lut.v[(c >> 8) & 255u] << 8) |
                 ((uint32_t)lut.v[(c >> 16) & 255u] << 16) | ((uint32_t)lut.v[c >> 24] << 24);
        }
        const uint64_t t = ((uint64_t)tz * dy + ty) * dx + tx;
        tex[2 * t] = w0;
        tex[2 * t + 1] = w1;
        texel[i] = (uint32_t)t;
    }
}

}  // namespace

namespace voxmi {

// Brick upload: one workgroup per request copies a (B+2)^3 brick from the
// voxelizer result into its pool slot (x-fastest rows of B+2 texels) and marks
// the node resident with the slot's texel pointer (BrickPool::uploadBrick,
// GigaVoxelsBrickPool.cpp 729-790).  Word 1 is written, not OR-ed: the
// reference ORs the new pointer into the old one (setBrickPtr), which only
// holds while a node is never uploaded twice.
__global__ void k_brick_upload(BrickSrc src, const uint4* __restrict__ req,
                               const uint32_t* __restrict__ nreq, uint32_t ex, uint32_t ey,
                               uint32_t ez, uint32_t sx, uint32_t sy, uint32_t dx, uint32_t dy,
                               uint32_t wide, uint32_t* __restrict__ col,
                               uint32_t* __restrict__ nrm, uint32_t* __restrict__ tex) {
    const uint32_t count = *nreq;
    const uint32_t nv = ex * ey * ez;
    for (uint32_t r = blockIdx.x; r < count; r += gridDim.x) {
        const uint4 q = req[r];  // depth, brick, slot, node texel
        const uint32_t* sc = src.col[q.x] + (uint64_t)q.y * nv;
        const uint32_t* sn = src.nrm[q.x] + (uint64_t)q.y * nv;
        const uint32_t ox = (q.z % sx) * ex;
        const uint32_t oy = ((q.z / sx) % sy) * ey;
        const uint32_t oz = (q.z / (sx * sy)) * ez;
        for (uint32_t v = threadIdx.x; v < nv; v += blockDim.x) {
            const uint32_t x = v % ex;
            const uint32_t yz = v / ex;
            const uint32_t y = yz % ey;
            const uint32_t z = yz / ey;
            const uint64_t t = ((uint64_t)(oz + z) * dy + (oy + y)) * dx + (ox + x);
            col[t] = sc[v];
            if (nrm) nrm[t] = sn[v];   // a colour-only pool keeps no normals
        }
        if (threadIdx.x == 0 && q.w != kNoTexel) {
            tex[2 * (uint64_t)q.w] |= kResidentBit;
            tex[2 * (uint64_t)q.w + 1] = wide ? q.z : pack10(ox + 1, oy + 1, oz + 1);
        }
    }
}

void launch_brick_upload(hipStream_t st, const BrickSrc& src, const uint4* req,
                         const uint32_t* nreq, uint32_t maxReq, const uint32_t e[3],
                         const uint32_t slots[3], const uint32_t dims[3], uint32_t wide,
                         uint32_t* col, uint32_t* nrm, uint32_t* tex) {
    if (maxReq == 0) return;
    const unsigned grid = maxReq < 4096u ? maxReq : 4096u;
    k_brick_upload<<<grid, kPoolBlock, 0, st>>>(src, req, nreq, e[0], e[1], e[2], slots[0],
                                                slots[1], dims[0], dims[1], wide, col, nrm, tex);
}

}  // namespace voxmi

// ---------------------------------------------------------------------------
// host side

struct vox_brick_pool {
    uint32_t B[3] = {0, 0, 0}, slots[3] = {0, 0, 0}, dims[3] = {0, 0, 0};
    uint32_t nslots = 0, resident = 0;
    bool valid = false;
    bool wide = false;   // vox_brick_pool_create_wide
    hipStream_t st = nullptr;
    DevBuf col, nrm;
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

struct vox_octree {
    vox_brick_pool* bp = nullptr;
    int L = 0;
    bool valid = false;
    uint32_t flags = 0;
    uint32_t ndim[3] = {0, 0, 0};
    uint64_t blocks = 0;
    DevBuf types[VOX_MAX_LEVELS], cc[VOX_MAX_LEVELS], bidx[VOX_MAX_LEVELS];
    DevBuf state[VOX_MAX_LEVELS], sub[VOX_MAX_LEVELS], pre[VOX_MAX_LEVELS];
    DevBuf texel[VOX_MAX_LEVELS];
    DevBuf tex, req, nreq;
    // Recorded after every write of the node texture (the build here, a
    // brick server's frame): readers of the texture on other streams wait
    // on it.
    hipEvent_t ready = nullptr;
    BrickSrc src{};
    uint64_t bricks[VOX_MAX_LEVELS] = {};
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

namespace {

unsigned grid_for(uint64_t n) {
    const uint64_t g = (n + kPoolBlock - 1) / kPoolBlock;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

// The constant colour a reader gets back from tree.gvx: the writer prints
// float(c) / 255.0 with the default 6 significant digits (Voxelizer.cu
// 2896-2914), the reader parses it into a float (QString::toFloat: double,
// then narrowed) and truncates value * 255.0f (GigaVoxelsOctTree.cpp
// 1615-1618).
ConstLut text_round_trip_lut() {
    ConstLut l;
    for (int c = 0; c < 256; ++c) {
        std::ostringstream os;
        os << static_cast<float>(c) / 255.0;
        const float f = (float)std::strtod(os.str().c_str(), nullptr);
        l.v[c] = static_cast<unsigned char>(f * 255.0f);
    }
    return l;
}

// OctTreeNodePool::create3DTexture's sizing (GigaVoxelsOctTreeNodePool.cpp
// 101-116): floor(pow(n, 1/3.0)) in double, made even (2 when it reaches 0),
// then x grown by 2 until the texture holds all nodes.
void texture_dims(uint64_t numNodes, uint32_t d[3]) {
    size_t cubeRoot = static_cast<size_t>(std::pow(static_cast<double>(numNodes), 1.0 / 3.0));
    if (cubeRoot % 2 != 0) {
        --cubeRoot;
        if (cubeRoot == 0) cubeRoot = 2;
    }
    size_t x = cubeRoot, y = cubeRoot, z = cubeRoot;
    while (x * y * z < numNodes) x += 2;
    d[0] = (uint32_t)x;
    d[1] = (uint32_t)y;
    d[2] = (uint32_t)z;
}

bool build_tree(vox_octree* t, const vox_result* r) {
    vox_brick_pool* bp = t->bp;
    hipStream_t st = bp->st;
    t->L = r->levels;
    if (t->L < 1 || t->L > VOX_MAX_LEVELS) return t->fail("octree: bad level count");
    for (int a = 0; a < 3; ++a)
        if (r->brick_dim[a] != bp->B[a])
            return t->fail("octree: brick dimensions differ from the brick pool's");
    // copy the node arrays of the result
    for (int d = 0; d < t->L; ++d) {
        const uint64_t n = r->nodes[d];
        if (n != (1ull << (3 * d))) return t->fail("octree: unexpected node count");
        if (!t->types[d].ensure(n, t->err, "types") || !t->cc[d].ensure(4 * n, t->err, "colors") ||
            !t->bidx[d].ensure(4 * n, t->err, "brick index") ||
            !t->state[d].ensure(n, t->err, "state") || !t->sub[d].ensure(4 * n, t->err, "sub") ||
            !t->pre[d].ensure(4 * n, t->err, "pre") || !t->texel[d].ensure(4 * n, t->err, "texel"))
            return false;
        if (!t->check(hipMemcpyAsync(t->types[d].p, r->types[d], n, hipMemcpyDeviceToDevice, st),
                      "copy types") ||
            !t->check(hipMemcpyAsync(t->cc[d].p, r->const_colors[d], 4 * n,
                                     hipMemcpyDeviceToDevice, st),
                      "copy colors") ||
            !t->check(hipMemcpyAsync(t->bidx[d].p, r->brick_index[d], 4 * n,
                                     hipMemcpyDeviceToDevice, st),
                      "copy brick index"))
            return false;
        t->src.col[d] = r->brick_colors[d];
        t->src.nrm[d] = r->brick_normals[d];
        t->bricks[d] = r->bricks[d];
        k_missing_bricks<<<grid_for(n), kPoolBlock, 0, st>>>(n, r->bricks[d], t->types[d].as<uint8_t>(),
                                                            t->bidx[d].as<uint32_t>());
    }
    // tree membership, top-down
    for (int d = 0; d < t->L; ++d) {
        const uint64_t n = 1ull << (3 * d);
        k_tree_state<<<grid_for(n), kPoolBlock, 0, st>>>(
            n, (uint32_t)d, d == t->L - 1 ? 1 : 0, t->types[d].as<uint8_t>(),
            d ? t->state[d - 1].as<uint8_t>() : nullptr, t->state[d].as<uint8_t>());
    }
    // subtree counts, bottom-up
    for (int d = t->L - 1; d >= 0; --d) {
        const uint64_t n = 1ull << (3 * d);
        k_tree_subtree<<<grid_for(n), kPoolBlock, 0, st>>>(
            n, (uint32_t)d, t->state[d].as<uint8_t>(),
            d + 1 < t->L ? t->sub[d + 1].as<uint32_t>() : nullptr, t->sub[d].as<uint32_t>());
    }
    uint32_t rootSub = 0;
    if (!t->check(hipMemcpyAsync(&rootSub, t->sub[0].p, 4, hipMemcpyDeviceToHost, st),
                  "read block count") ||
        !t->check(hipStreamSynchronize(st), "synchronize"))
        return false;
    t->blocks = 1ull + rootSub;
    texture_dims(8 * t->blocks, t->ndim);
    for (int a = 0; a < 3; ++a)
        if (t->ndim[a] > 1024)
            return t->fail("octree: node texture exceeds 1024 texels per axis (10-bit pointers)");
    const uint64_t ntex = (uint64_t)t->ndim[0] * t->ndim[1] * t->ndim[2];
    if (!t->tex.ensure(8 * ntex, t->err, "node texture")) return false;
    launch_fill_u32(st, t->tex.as<uint32_t>(), 2 * ntex, 0u);
    // preorder ranks, top-down (the root has rank 0)
    launch_fill_u32(st, t->pre[0].as<uint32_t>(), 1, 0u);
    for (int d = 0; d + 1 < t->L; ++d) {
        const uint64_t n = 1ull << (3 * d);
        k_tree_rank<<<grid_for(n), kPoolBlock, 0, st>>>(
            n, (uint32_t)d, t->state[d].as<uint8_t>(), t->pre[d].as<uint32_t>(),
            t->sub[d + 1].as<uint32_t>(), t->pre[d + 1].as<uint32_t>());
    }
    const uint32_t bx = t->ndim[0] / 2, by = t->ndim[1] / 2;
    ConstLut lut;
    if (t->flags & VOX_TREE_EXACT_CONST) {
        for (int c = 0; c < 256; ++c) lut.v[c] = (uint8_t)c;
    } else {
        lut = text_round_trip_lut();
    }
    const int storeConst = (t->flags & VOX_TREE_REFERENCE_CONST) ? 0 : 1;
    for (int d = 0; d < t->L; ++d) {
        const uint64_t n = 1ull << (3 * d);
        k_tree_write<<<grid_for(n), kPoolBlock, 0, st>>>(
            n, (uint32_t)d, t->state[d].as<uint8_t>(), t->types[d].as<uint8_t>(),
            t->cc[d].as<uint32_t>(), t->pre[d].as<uint32_t>(),
            d ? t->pre[d - 1].as<uint32_t>() : nullptr, bx, by, t->ndim[0], t->ndim[1], lut,
            storeConst, t->tex.as<uint32_t>(), t->texel[d].as<uint32_t>());
    }
    if (!t->check(hipGetLastError(), "node pool kernels")) return false;
    // uploadInitialBricks: BrickPool::initBricks(root, 1) puts a NON-CONST
    // root's brick into the next slot (slot 0 for the first octree).
    uint8_t rootType = 0;
    if (!t->check(hipMemcpyAsync(&rootType, t->types[0].p, 1, hipMemcpyDeviceToHost, st),
                  "read root type") ||
        !t->check(hipStreamSynchronize(st), "synchronize"))
        return false;
    // host staging of the one request; both stay alive until the final sync
    const uint4 q = make_uint4(0u, 0u, bp->resident, 0u);  // root texel is 0
    const uint32_t one = 1;
    if (rootType == 1 && t->bricks[0] > 0 && bp->resident < bp->nslots &&
        !(t->flags & VOX_TREE_NO_ROOT_BRICK)) {
        if (!t->req.ensure(sizeof(uint4), t->err, "request") ||
            !t->nreq.ensure(4, t->err, "request count"))
            return false;
        if (!t->check(hipMemcpyAsync(t->req.p, &q, sizeof(q), hipMemcpyHostToDevice, st), "h2d") ||
            !t->check(hipMemcpyAsync(t->nreq.p, &one, 4, hipMemcpyHostToDevice, st), "h2d"))
            return false;
        const uint32_t e[3] = {bp->B[0] + 2, bp->B[1] + 2, bp->B[2] + 2};
        launch_brick_upload(st, t->src, t->req.as<uint4>(), t->nreq.as<uint32_t>(), 1, e,
                            bp->slots, bp->dims, bp->wide ? 1u : 0u, bp->col.as<uint32_t>(),
                            bp->nrm.as<uint32_t>(), t->tex.as<uint32_t>());
        ++bp->resident;
    }
    return t->check(hipGetLastError(), "brick upload") &&
           t->check(hipEventRecord(t->ready, st), "event record") &&
           t->check(hipStreamSynchronize(st), "synchronize");
}

}  // namespace

namespace voxmi {

bool octree_refs(vox_octree* t, OctreeRefs& r) {
    if (!t || !t->valid || !t->bp || !t->bp->valid) return false;
    for (int a = 0; a < 3; ++a) {
        r.ndim[a] = t->ndim[a];
        r.B[a] = t->bp->B[a];
        r.poolDims[a] = t->bp->dims[a];
        r.poolSlots[a] = t->bp->slots[a];
    }
    r.poolWide = t->bp->wide ? 1u : 0u;
    r.tex = t->tex.as<uint32_t>();
    r.levels = t->L;
    r.poolCol = t->bp->col.as<uint32_t>();
    r.poolNrm = t->bp->nrm.as<uint32_t>();
    r.st = t->bp->st;
    r.ready = t->ready;
    r.pool = t->bp;
    for (int d = 0; d < kMaxLevels; ++d) {
        const bool in = d < t->L;
        r.types[d] = in ? t->types[d].as<uint8_t>() : nullptr;
        r.bidx[d] = in ? t->bidx[d].as<uint32_t>() : nullptr;
        r.texel[d] = in ? t->texel[d].as<uint32_t>() : nullptr;
    }
    r.src = t->src;
    return true;
}

bool pool_refs(vox_brick_pool* bp, PoolRefs& r) {
    if (!bp || !bp->valid) return false;
    for (int a = 0; a < 3; ++a) {
        r.B[a] = bp->B[a];
        r.dims[a] = bp->dims[a];
        r.slots[a] = bp->slots[a];
    }
    r.wide = bp->wide ? 1u : 0u;
    r.col = bp->col.as<uint32_t>();
    r.nrm = bp->nrm.as<uint32_t>();
    r.st = bp->st;
    return true;
}

bool octree_record_written(vox_octree* t, hipStream_t st) {
    return t && t->ready && hipEventRecord(t->ready, st) == hipSuccess;
}

}  // namespace voxmi

extern "C" {

vox_brick_pool* vox_brick_pool_create(const uint32_t brick_dim[3], const uint32_t slots[3]) {
    vox_brick_pool* bp = new (std::nothrow) vox_brick_pool;
    if (!bp) return nullptr;
    if (!brick_dim) {
        bp->fail("brick pool: brick_dim is NULL");
        return bp;
    }
    const bool wide = brick_dim[0] > 14 || brick_dim[1] > 14 || brick_dim[2] > 14;
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        bp->B[a] = brick_dim[a];
        bp->slots[a] = slots ? slots[a] : (wide ? 8u : 16u);
        bp->dims[a] = (brick_dim[a] + 2) * bp->slots[a];
        total *= bp->dims[a];
        if (brick_dim[a] == 0 || bp->slots[a] == 0) {
            bp->fail("brick pool: zero brick or slot dimension");
            return bp;
        }
        if (bp->dims[a] > 1024) {
            bp->fail("brick pool: more than 1024 texels per axis do not fit 10-bit brick pointers");
            return bp;
        }
    }
    bp->nslots = bp->slots[0] * bp->slots[1] * bp->slots[2];
    if (!bp->check(hipStreamCreateWithFlags(&bp->st, hipStreamNonBlocking), "hipStreamCreate"))
        return bp;
    if (!bp->col.ensure(4 * total, bp->err, "brick pool colors") ||
        !bp->nrm.ensure(4 * total, bp->err, "brick pool normals"))
        return bp;
    // The reference creates the textures without data; cleared here.
    launch_fill_u32(bp->st, bp->col.as<uint32_t>(), total, 0u);
    launch_fill_u32(bp->st, bp->nrm.as<uint32_t>(), total, 0u);
    if (!bp->check(hipStreamSynchronize(bp->st), "synchronize")) return bp;
    bp->valid = true;
    return bp;
}

namespace {

vox_brick_pool* create_wide(const uint32_t brick_dim[3], uint64_t nslots, bool normals) {
    vox_brick_pool* bp = new (std::nothrow) vox_brick_pool;
    if (!bp) return nullptr;
    if (!brick_dim) {
        bp->fail("brick pool: brick_dim is NULL");
        return bp;
    }
    if (nslots == 0 || nslots > (1ull << 30)) {
        bp->fail("wide brick pool: between 1 and 2^30 slots");
        return bp;
    }
    uint64_t e3 = 1;
    for (int a = 0; a < 3; ++a) {
        if (brick_dim[a] == 0) {
            bp->fail("brick pool: zero brick dimension");
            return bp;
        }
        bp->B[a] = brick_dim[a];
        e3 *= brick_dim[a] + 2;
    }
    // slots stacked along z: slot s holds texels [s * e^3, (s + 1) * e^3)
    if ((uint64_t)(brick_dim[2] + 2) * nslots >= (1ull << 32)) {
        bp->fail("wide brick pool: slots * (B+2) above 2^32 texels along z");
        return bp;
    }
    bp->wide = true;
    bp->slots[0] = bp->slots[1] = 1;
    bp->slots[2] = (uint32_t)nslots;
    bp->dims[0] = brick_dim[0] + 2;
    bp->dims[1] = brick_dim[1] + 2;
    bp->dims[2] = (brick_dim[2] + 2) * (uint32_t)nslots;
    bp->nslots = (uint32_t)nslots;
    const uint64_t total = e3 * nslots;
    if (!bp->check(hipStreamCreateWithFlags(&bp->st, hipStreamNonBlocking), "hipStreamCreate"))
        return bp;
    if (!bp->col.ensure(4 * total, bp->err, "brick pool colors") ||
        (normals && !bp->nrm.ensure(4 * total, bp->err, "brick pool normals")))
        return bp;
    launch_fill_u32(bp->st, bp->col.as<uint32_t>(), total, 0u);
    if (normals) launch_fill_u32(bp->st, bp->nrm.as<uint32_t>(), total, 0u);
    if (!bp->check(hipStreamSynchronize(bp->st), "synchronize")) return bp;
    bp->valid = true;
    return bp;
}

}  // namespace

vox_brick_pool* vox_brick_pool_create_wide(const uint32_t brick_dim[3], uint64_t nslots) {
    return create_wide(brick_dim, nslots, true);
}

vox_brick_pool* vox_brick_pool_create_colour(const uint32_t brick_dim[3], uint64_t nslots) {
    return create_wide(brick_dim, nslots, false);
}

void vox_brick_pool_destroy(vox_brick_pool* bp) {
    if (!bp) return;
    if (bp->st) (void)hipStreamSynchronize(bp->st);
    bp->col.release();
    bp->nrm.release();
    if (bp->st) (void)hipStreamDestroy(bp->st);
    delete bp;
}

const char* vox_brick_pool_get_error(vox_brick_pool* bp) {
    if (!bp) return "null brick pool";
    bp->errCache = bp->err.str();
    return bp->errCache.c_str();
}

int vox_brick_pool_get_view(vox_brick_pool* bp, vox_brick_pool_view* out) {
    if (!bp || !bp->valid || !out) return 0;
    for (int a = 0; a < 3; ++a) {
        out->brick_dim[a] = bp->B[a];
        out->slots[a] = bp->slots[a];
        out->dims[a] = bp->dims[a];
    }
    out->colors = bp->col.as<uint32_t>();
    out->normals = bp->nrm.as<uint32_t>();
    out->resident = bp->resident;
    out->wide = bp->wide ? 1u : 0u;
    return 1;
}

vox_octree* vox_octree_create(vox_brick_pool* bp, const vox_result* tree, uint32_t flags) {
    vox_octree* t = new (std::nothrow) vox_octree;
    if (!t) return nullptr;
    t->bp = bp;
    t->flags = flags;
    if (!bp || !bp->valid) {
        t->fail("octree: invalid brick pool");
        return t;
    }
    if (!tree) {
        t->fail("octree: tree is NULL");
        return t;
    }
    if (flags & ~(VOX_TREE_REFERENCE_CONST | VOX_TREE_EXACT_CONST | VOX_TREE_NO_ROOT_BRICK)) {
        t->fail("octree: unknown flags");
        return t;
    }
    if (!t->check(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming), "hipEventCreate"))
        return t;
    t->valid = build_tree(t, tree);
    return t;
}

void vox_octree_destroy(vox_octree* t) {
    if (!t) return;
    if (t->bp && t->bp->st) (void)hipStreamSynchronize(t->bp->st);
    for (int d = 0; d < VOX_MAX_LEVELS; ++d) {
        t->types[d].release();
        t->cc[d].release();
        t->bidx[d].release();
        t->state[d].release();
        t->sub[d].release();
        t->pre[d].release();
        t->texel[d].release();
    }
    t->tex.release();
    t->req.release();
    t->nreq.release();
    if (t->ready) (void)hipEventDestroy(t->ready);
    delete t;
}

const char* vox_octree_get_error(vox_octree* t) {
    if (!t) return "null octree";
    t->errCache = t->err.str();
    return t->errCache.c_str();
}

int vox_octree_get_view(vox_octree* t, vox_octree_view* out) {
    if (!t || !t->valid || !out) return 0;
    std::memset(out, 0, sizeof(*out));
    out->levels = t->L;
    for (int a = 0; a < 3; ++a) out->node_dim[a] = t->ndim[a];
    out->node_blocks = t->blocks;
    out->nodes = t->tex.as<uint32_t>();
    for (int d = 0; d < t->L; ++d) out->node_texel[d] = t->texel[d].as<uint32_t>();
    return 1;
}

int vox_pool_copy_to_host(void* dst, const void* src_device, uint64_t nbytes) {
    if (!dst || !src_device) return 0;
    // the library's streams are non-blocking: wait for all of them
    if (hipDeviceSynchronize() != hipSuccess) return 0;
    return hipMemcpy(dst, src_device, nbytes, hipMemcpyDeviceToHost) == hipSuccess ? 1 : 0;
}

}  // extern "C"
