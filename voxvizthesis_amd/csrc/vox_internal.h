// vox_internal.h -- shared declarations between the HIP kernels
// (vox_kernels.hip) and the host orchestration (vox_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <sstream>

#include "../../include/voxmi_pool.h"
#include "../../include/voxmi_serve.h"

namespace voxmi {

// Device blocks handed back by recycle(): kept in a process-wide cache and
// handed out again by ensure() instead of a new hipMalloc, so that buffers a
// handle releases do not cost a hipFree (which waits for the whole device).
// block_take returns a cached block of at least `bytes` and at most a quarter
// more (its size in *cap), or null; block_flush frees every cached block.
void* block_take(size_t bytes, size_t* cap);
void block_give(void* p, size_t cap);
void block_flush();

// Grow-only device buffer; allocation failures are reported into err.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes, std::stringstream& err, const char* what) {
        if (bytes <= cap && p) return true;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        size_t b = std::max<size_t>(bytes, 256);
        size_t got = 0;
        if ((p = block_take(b, &got)) != nullptr) {
            cap = got;
            return true;
        }
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            // cached blocks may hold the memory: free them and try again
            (void)hipGetLastError();
            block_flush();
            e = hipMalloc(&p, b);
        }
        if (e != hipSuccess) {
            err << "hipMalloc failed for " << what << " (" << b << " bytes): "
                << hipGetErrorString(e) << std::endl;
            // the runtime keeps the failure as its last error: clear it, or
            // the next launch check would report it again
            (void)hipGetLastError();
            p = nullptr;
            return false;
        }
        cap = b;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    // Give the block to the cache instead of freeing it.  Only when no
    // queued device work can still use it (its stream synchronised, and no
    // other stream reads it).
    void recycle() {
        if (p) block_give(p, cap);
        p = nullptr;
        cap = 0;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

constexpr int kTile = 8;            // voxel tile edge (8^3 = 512 voxels)
constexpr int kTileVox = 512;
constexpr int kTileWords = 16;      // 512 bits per tile
constexpr int kBlock = 256;         // threads per workgroup for tile kernels
constexpr int kFillDepth = 19;      // terrain fill reach (VoxelizerKernels.cu 888)
// Hit records of the voxelize pass go to kRecSegs segments, unit u of the
// pass (kVoxTiles consecutive tiles of its work list, at most kVoxTiles *
// 512 records) appending to segment u % kRecSegs with the segment's own
// counter (no single hot counter).  A segment thus takes at most
// ceil(units / kRecSegs) units, and the capacity below (sized as if
// workgroup b of a grid-stride launch of `grid` workgroups filled segment
// b % kRecSegs) is never smaller: ceil(units / grid) * ceil(grid / kRecSegs)
// >= ceil(units / kRecSegs), whichever workgroup runs a unit.
constexpr uint32_t kRecSegs = 64;
constexpr int kVoxTiles = 2;
inline uint64_t rec_seg_cap(uint64_t nActive, uint64_t grid) {
    if (grid == 0) return 0;
    const uint64_t units = (nActive + kVoxTiles - 1) / kVoxTiles;
    return ((units + grid - 1) / grid) * kVoxTiles * ((grid + kRecSegs - 1) / kRecSegs) * 512ull;
}
// workgroups of the voxelize launch over nActive tiles
inline uint64_t vox_grid(uint64_t nActive, uint64_t persistent) {
    const uint64_t units = (nActive + kVoxTiles - 1) / kVoxTiles;
    return units < persistent ? units : persistent;
}
constexpr int kFillTiles = 3;       // ceil(19 / 8)

struct VoxParams {
    uint32_t S[3];        // voxel space per axis (N + 4*extra)
    uint32_t TG[3];       // level-0 tile grid, ceil(S/8)
    uint64_t numTiles;    // TG[0]*TG[1]*TG[2]
    float offP[3];        // p - extra*delta
    float dP[3];          // delta
    // Chunk edge per axis (the reference's voxChunkDim; S when one chunk
    // holds the voxel space): the colour pass measures voxel centres from
    // the chunk's origin and the terrain fill stops at the chunk's z = 0.
    uint32_t C[3];
};

struct GroupArgs {
    uint32_t triOff;      // first global triangle index of the group
    uint32_t ntri;
    // voxels the group's triangles may reach: z from the group's bounds, x
    // and y the chunks its bounds overlap (the whole space with one chunk)
    uint32_t zStart, zEnd;
    uint32_t xStart, xEnd, yStart, yEnd;
    int32_t overlaps;     // group bbox overlaps the chunk
    int32_t textured;
    int32_t terrain;
    int32_t hasVN;
    const float* vn;      // group-local vertex normals (9 per triangle) or null
    const float* uv;      // group-local uvs (6 per triangle) or null
    const uint8_t* img;   // RGBA8 image or null
    int32_t w, h, addr0, addr1;
};

struct LevelMap {         // sparse tiles of one mip level
    uint32_t dims[3];     // S_k
    uint32_t TG[3];       // tile grid
    uint64_t numTiles;
    const int32_t* tileIdx;   // dense map tile -> compact index or -1
    // Level 0 stores only the voxels that some group hit (defOcc, the
    // occupancy words) or a terrain fill reached (defFill), 16 words per
    // active tile; any other voxel of an active tile reads as zero colour,
    // zero normal.  Both null for the mip levels, which are written in full.
    const uint32_t* defOcc = nullptr;
    const uint32_t* defFill = nullptr;
};

// Brick arrays of one octree, per depth (the voxelizer result's bricks).
constexpr int kMaxLevels = 16;
struct BrickSrc {
    const uint32_t* col[kMaxLevels];
    const uint32_t* nrm[kMaxLevels];
};

// What the ray cast and the brick server need of an octree and its pool.
struct OctreeRefs {
    uint32_t ndim[3];       // node texture texels per axis
    uint32_t* tex;          // node texture (2 words per texel)
    int32_t levels;
    uint32_t B[3];          // brick interior
    uint32_t poolDims[3];   // brick pool texels per axis
    uint32_t poolSlots[3];
    // wide pool: slots stacked along z, word 1 of a resident node holds the
    // slot index instead of 10-bit texel coordinates (no 1024-texel limit)
    uint32_t poolWide;
    uint32_t* poolCol;
    uint32_t* poolNrm;
    hipStream_t st;         // the pool's stream
    hipEvent_t ready;       // recorded after the last write of the node texture
    vox_brick_pool* pool;
    // per depth: node types, brick index, node texel (0xffffffff: not in tree)
    const uint8_t* types[kMaxLevels];
    const uint32_t* bidx[kMaxLevels];
    const uint32_t* texel[kMaxLevels];
    BrickSrc src;           // the voxelizer result's bricks
};
bool octree_refs(vox_octree* t, OctreeRefs& r);

// The request list of a ray caster's last frame, all on the device: distinct
// node ids, the device clock when each was appended, the count (the first
// of the ray caster's counters, at most maxIds), the event recorded when
// the list was complete, and the tree it was cast against.
struct RequestRefs {
    const uint32_t* ids;
    const uint64_t* times;
    const uint32_t* count;
    uint32_t maxIds;
    hipEvent_t done;
    vox_octree* tree;
};
bool raycaster_requests(vox_raycaster* rc, RequestRefs& r);

// A brick pool's storage and its stream.
struct PoolRefs {
    uint32_t B[3], dims[3], slots[3];
    uint32_t wide;          // see OctreeRefs::poolWide
    uint32_t* col;
    uint32_t* nrm;
    hipStream_t st;
};
bool pool_refs(vox_brick_pool* bp, PoolRefs& r);
// Record the octree's ready event on `st` after work there wrote its node
// texture.
bool octree_record_written(vox_octree* t, hipStream_t st);

// ---- launchers (all asynchronous on `st`) ----
// Up to kGroupSet groups in one launch: group i (index g[i] of the handle)
// takes the workgroups start[i] .. start[i + 1] - 1.
constexpr int kGroupSet = 8;
struct GroupSet {
    GroupArgs G[kGroupSet];
    uint32_t start[kGroupSet + 1];
    int32_t g[kGroupSet];
    int n;
};
// Setup (bounds, face normal, bin counts) and scatter of every group, a
// launch per kGroupSet groups.
void launch_setup(hipStream_t st, const VoxParams& P, const std::vector<GroupArgs>& groups,
                  const float* verts, float* fnorm, uint32_t* bbox, uint32_t* binCount);
void launch_scatter(hipStream_t st, const VoxParams& P, const std::vector<GroupArgs>& groups,
                    const uint32_t* bbox, const uint32_t* binStart, uint32_t* binFill,
                    uint32_t* binTri);
void launch_activity0(hipStream_t st, const VoxParams& P, int ngroups, const int32_t* grpFill,
                      const uint32_t* binStart, uint32_t* mark);
// Also writes excl[n] (the number of marked tiles) to *count.
void launch_compact(hipStream_t st, uint64_t n, const uint32_t* mark, const uint32_t* excl,
                    int32_t* tileIdx, int32_t* list, uint32_t* count);
// mark[t] = 1 when any of the (up to 8) source tiles under level tile t is
// active, else 0: every mark written, no clearing first.
void launch_mark_children(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                          uint32_t* markDst);
void launch_voxelize(hipStream_t st, const VoxParams& P, const GroupArgs& G, int g, int init,
                     uint64_t nActive, const int32_t* list, const float* verts,
                     const float* fnorm, const uint32_t* bbox, const uint32_t* binStart,
                     const uint32_t* binTri, uint32_t* occ, uint32_t* hit,
                     unsigned long long* satCounter, uint4* work, uint32_t* workCount,
                     uint4* rec, uint32_t* recCount, unsigned persistentBlocks);
// One job of a batched voxelize launch: the arguments of one group's
// k_voxelize launch (its work list and count written on the device by that
// job's group_tiles pass).
struct VoxUnitJob {
    VoxParams P;
    GroupArgs G;
    int32_t init;
    const uint4* work;
    const uint32_t* workCount;
    const float* verts;
    const float* fnorm;
    const uint32_t* bbox;
    const uint32_t* binTri;
    uint32_t* occ;
    uint32_t* hit;
    uint4* rec;
    uint32_t* recCount;
    uint64_t segCap;
};
// The arguments launch_voxelize would pass for one job (nothing launched).
VoxUnitJob voxelize_job(const VoxParams& P, const GroupArgs& G, int init, uint64_t nActive,
                        const float* verts, const float* fnorm, const uint32_t* bbox,
                        const uint32_t* binTri, uint32_t* occ, uint32_t* hit, uint4* work,
                        uint32_t* workCount, uint4* rec, uint32_t* recCount,
                        unsigned persistentBlocks);
// The voxelize pass of njobs jobs (device table) in one persistent launch of
// `blocks` workgroups over a device queue; unitStart holds njobs + 1 words,
// queue one.
void launch_voxelize_batch(hipStream_t st, const VoxUnitJob* jobs, uint32_t njobs,
                           uint32_t* unitStart, uint32_t* queue, unsigned blocks);
void launch_group_tiles(hipStream_t st, const VoxParams& P, int g, uint64_t nActive,
                        const int32_t* list, const uint32_t* binStart, uint4* work,
                        uint32_t* workCount, uint32_t* recCount);
// fillw: level 0's filled-voxel words (a voxel neither hit before this group
// nor filled holds no stored value and counts as zero).
void launch_colour(hipStream_t st, const VoxParams& P, const GroupArgs& G, int init,
                   uint64_t nActive, const uint4* rec, const uint32_t* recCount,
                   const int32_t* list, const float* verts, const float* fnorm, uint32_t* col,
                   uint32_t* qn, float* nrm, const uint32_t* fillw, unsigned persistentBlocks);
// Also records the voxels it fills in fillw (16 words per active tile; set
// on the first group, OR-ed after).
void launch_fill(hipStream_t st, const VoxParams& P, int g, int init, uint64_t nActive,
                 const int32_t* list, const int32_t* tileIdx, const uint32_t* binStart,
                 uint32_t* col, uint32_t* qn, float* nrm, const uint32_t* hit, uint32_t* fillw);
// Stored source voxels the mip filter of dst reads (active source tiles
// under its active tiles; level 0: hit or filled voxels), added to *out.
void launch_mip_source_count(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                             uint64_t nActive, const int32_t* listDst, unsigned long long* out);
void launch_mip(hipStream_t st, const LevelMap& src, const LevelMap& dst, uint64_t nActive,
                const int32_t* listDst, const uint32_t* colSrc, const float* nrmSrc,
                uint32_t* colDst, uint32_t* qnDst, float* nrmDst);
// Node samples of every depth in one launch (node typing, first half):
// depth d's nodes are the threads start[d] <= n < start[d + 1]; cc and flag
// of depth d get each node's sample and whether its brick box reaches an
// inactive tile.
struct SampleDepths {
    LevelMap lm[kMaxLevels];
    const uint32_t* col[kMaxLevels];
    uint32_t* cc[kMaxLevels];
    uint32_t* flag[kMaxLevels];
    uint32_t Rd[kMaxLevels];
    uint32_t border[kMaxLevels][3];
    uint32_t xLimit[kMaxLevels];
    uint64_t start[kMaxLevels + 1];
    uint32_t B[3];
    int L;
};
void launch_node_sample_all(hipStream_t st, const SampleDepths& S);
// Node typing, second half: every voxel of the active tiles compared with the
// samples of the nodes whose bricks hold it.
void launch_node_compare(hipStream_t st, const LevelMap& lm, uint64_t nActive, const int32_t* list,
                         const uint32_t* col, uint32_t Rd, const uint32_t B[3],
                         const uint32_t border[3], uint32_t xLimit, uint32_t* cc, uint32_t* flag);
// Brick indices of every depth in one launch over the all-depth node array:
// node n of depth d (off[d] <= n < off[d + 1]) gets type, brick index
// (excl[n] - excl[off[d]]) and its brick's node.
struct DepthOut {
    uint32_t* bidx[kMaxLevels];
    uint32_t* brickNode[kMaxLevels];
    uint8_t* types[kMaxLevels];
    uint64_t off[kMaxLevels + 1];
    int L;
};
void launch_brick_assign_all(hipStream_t st, const uint32_t* flag, const uint32_t* excl,
                             const DepthOut& D);
// Range boundaries of one scan, at most kMaxLevels ranges (passed by value).
struct RangeOffs {
    uint64_t v[kMaxLevels + 1];
};
// out[d] = excl[offs[d + 1]] - excl[offs[d]] for d < n (n <= kMaxLevels).
void launch_range_counts(hipStream_t st, const uint32_t* excl, const uint64_t* offs, int n,
                         uint32_t* out);
// The bricks of every depth in one launch: entry e (a depth with bricks)
// covers workgroups start[e] <= w < start[e + 1], one brick each.
struct GatherDepths {
    LevelMap lm[kMaxLevels];
    const uint32_t* col[kMaxLevels];
    const uint32_t* qn[kMaxLevels];
    const uint32_t* brickNode[kMaxLevels];
    uint32_t* bcol[kMaxLevels];
    uint32_t* bnrm[kMaxLevels];
    uint32_t Rd[kMaxLevels];
    uint32_t border[kMaxLevels][3];
    uint64_t start[kMaxLevels + 1];
    uint32_t B[3];
    int n;
};
void launch_brick_gather_all(hipStream_t st, const GatherDepths& G);
void launch_collapse(hipStream_t st, uint32_t Rd, int childIsLeaf, uint8_t* types,
                     const uint8_t* childTypes);

// ---- several chunks (generateVoxelsAndOctTree's chunk loop, Voxelizer.cu
// 2180-2814, with chunks smaller than the voxel space) ----
// Level k is laid out chunk by chunk: chunk i of an axis holds voxels
// [i * Ck, (i + 1) * Ck), Ck = C >> k, of that chunk's own mip chain.
constexpr int kMaxChunksAxis = 64;
struct ChunkGeom {
    uint32_t C[3];        // chunk edge at level 0
    uint32_t nc[3];       // chunks per axis (origins below N)
    // visit rank of chunk (i, j, l) at [(l * nc[1] + j) * nc[0] + i]
    const uint32_t* rank;
};
// mark[t] = 1 when a voxel of level tile t reads an active source tile
// (source voxels 2x, 2x + 1 of the same chunk, local coordinates).
// (ks: the source level, kd: the level built from it.)
void launch_chunk_mark(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                       const ChunkGeom& cg, int ks, int kd, uint32_t* markDst);
// The chunked mip filter as the single-chunk one with a per-chunk source
// offset (MipShift), where the chunk geometry allows it: every axis's chunk
// extent at the level a multiple of the tile edge, the per-chunk offset a
// multiple of two tiles and every source inside the source level.
struct MipShift {
    uint32_t cd[3];    // chunk extent at the level (0: one chunk, no shift)
    uint32_t off[3];   // source offset per chunk: Cs - 2 Cd
};
bool mip_shift_for(const LevelMap& src, const LevelMap& dst, const ChunkGeom& cg, int ks, int kd,
                   MipShift& out);
void launch_mark_children_shifted(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                                  const MipShift& shift, uint32_t* markDst);
void launch_mip_shifted(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                        const MipShift& shift, uint64_t nActive, const int32_t* listDst,
                        const uint32_t* colSrc, const float* nrmSrc, uint32_t* colDst,
                        uint32_t* qnDst, float* nrmDst);
void launch_chunk_mip(hipStream_t st, const LevelMap& src, const LevelMap& dst,
                      const ChunkGeom& cg, int ks, int kd, uint64_t nActive,
                      const int32_t* listDst, const uint32_t* colSrc, const float* nrmSrc,
                      uint32_t* colDst, uint32_t* qnDst, float* nrmDst);
// Node types and constant colours of every depth d of S (level L - 1 - d,
// S.lm / S.col; S.cc, S.flag; nodes numbered from S.start[d]): flag[n] = 1
// when a chunk the node's stored brick reaches holds a typed voxel of the
// brick unequal to that chunk's sample; cc[n] = the sample of the last chunk
// in visit order.
void launch_chunk_node_type(hipStream_t st, const SampleDepths& S, const ChunkGeom& cg);
// Brick keys of one depth: for each NON-CONST node whose stored brick lies
// in processed chunks, key = (rank of the chunk that completes it << (1 +
// 3d)) | (0 inside one chunk, 1 across chunks) << 3d | node; keys and nodes
// appended at *count.
void launch_chunk_brick_keys(hipStream_t st, const uint32_t* flag, const ChunkGeom& cg, int k,
                             uint32_t Rd, const uint32_t B[3], const uint32_t border[3],
                             unsigned long long* keys, uint32_t* nodes, uint32_t* count);
// Sort n (key, node) pairs by key (tmp: temporary storage; *tmpBytes is
// its size on entry, or the size needed when tmp is null).
hipError_t chunk_sort_pairs(hipStream_t st, void* tmp, size_t* tmpBytes,
                            const unsigned long long* keysIn, unsigned long long* keysOut,
                            const uint32_t* nodesIn, uint32_t* nodesOut, uint32_t n, int keyBits);
// After the sort: types[node] = flag, bidx[node] = 0xffffffff, then
// bidx[sorted[j]] = j and brickNode[j] = sorted[j].
void launch_chunk_brick_assign(hipStream_t st, uint64_t nodes, const uint32_t* flag,
                               const uint32_t* sorted, uint32_t nb, uint8_t* types,
                               uint32_t* bidx, uint32_t* brickNode);
// The whole bitmap (no clear needed) when S[0] is a multiple of 32 and a row
// of tile indices fits in LDS; false (nothing launched) otherwise.
bool launch_occupancy_rows(hipStream_t st, const VoxParams& P, const int32_t* tileIdx,
                           const uint32_t* occ, uint32_t* bitmap);
void launch_occupancy(hipStream_t st, const VoxParams& P, uint64_t nActive, const int32_t* list,
                      const uint32_t* occ, uint32_t* bitmap);
void launch_dense_level(hipStream_t st, const LevelMap& lm, const uint32_t* col,
                        const float* nrm, uint32_t* outCol, float* outNrm);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);
// Up to four (pointer, words, value) fills in one launch.
struct FillSeg {
    uint32_t* p;
    uint64_t n;
    uint32_t v;
};
void launch_fill_segs(hipStream_t st, const FillSeg* segs, int nseg);
// Copy n16 16-byte words.
void launch_copy_u128(hipStream_t st, const void* src, void* dst, uint64_t n16);
// nblocks 256-byte blocks (a power of two), block i of the launch copying
// block (i * 0x9E3779B1) mod nblocks (odd multiplier: a permutation)
void launch_copy_scattered(hipStream_t st, const void* src, void* dst, uint64_t nblocks);
// dst row k = src row idx[k] - base, w floats per row.
void launch_gather_rows(hipStream_t st, const float* src, const uint32_t* idx, uint64_t n,
                        uint32_t base, uint32_t w, float* dst);
// Vertex bounds of n triangles into out[6] (ordered ints: min xyz, max xyz;
// out must hold INT_MAX x3, INT_MIN x3 first).
void launch_tri_bounds(hipStream_t st, const float* v, uint64_t n, int32_t* out);
// One buffer of a multi-buffer hash: n words at p hashed with salt into
// outs[out]; block0 = the first block of the launch that works on it
// (segments in ascending block0, none empty).
struct HashSeg {
    const uint32_t* p;
    uint64_t n;
    uint64_t salt;
    uint32_t out;
    uint32_t block0;
};
void launch_hash_segs(hipStream_t st, const HashSeg* segs, uint32_t nseg, uint32_t blocks,
                      unsigned long long* outs);
// Block compression of bricks: BC3 colours (16 B per block) and BC1 normals
// (8 B per block), ceil(e/4)^2 * e blocks per brick of e^3 texels.
uint64_t bc_blocks_per_brick(const uint32_t e[3]);
void launch_bc_bricks(hipStream_t st, const uint32_t* col, const uint32_t* nrm, uint64_t nbricks,
                      const uint32_t e[3], uint32_t* outCol, uint32_t* outNrm);
// Upload requests (depth, brick, slot, node texel or 0xffffffff) into brick
// pool slots; *nreq requests, at most maxReq.
void launch_brick_upload(hipStream_t st, const BrickSrc& src, const uint4* req,
                         const uint32_t* nreq, uint32_t maxReq, const uint32_t e[3],
                         const uint32_t slots[3], const uint32_t dims[3], uint32_t wide,
                         uint32_t* col, uint32_t* nrm, uint32_t* tex);
// exclusive scan of n u32 values; out has n+1 entries, out[n] = total.
// tmp must hold scan_tmp_words(n) words, all zero before its first scan
// (ensure_scan_tmp allocates it so); a scan leaves it ready for the next.
// Scans that share tmp must not overlap in time; in and out are 16-byte
// aligned.  err (may be null): a device word that gets bit 0 set when a
// workgroup gave up waiting on a predecessor's published sum (a bounded wait
// that every schedule ends well within), after which out is not valid.
uint64_t scan_tmp_words(uint64_t n);
bool ensure_scan_tmp(DevBuf& b, uint64_t n, hipStream_t st, std::stringstream& err,
                     const char* what);
void launch_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                 uint32_t* err = nullptr);
// The same scan, leaving the input zeroed (for counters that count again).
void launch_scan_clear(hipStream_t st, uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                       uint32_t* err = nullptr);

}  // namespace voxmi
