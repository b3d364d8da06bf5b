// vox_api.cpp -- host orchestration and C ABI (include/voxmi.h).
//
// Mirrors cuda::Voxelizer (Voxelizer.h 22-244, Voxelizer.cu): the same
// call sequence and return conventions, but the whole pipeline runs on one
// HIP stream with everything resident in HBM.  Host round trips are limited
// to three small counter reads per vox_generate (bin entries, active tiles
// per level, bricks per level) that size the next allocations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <new>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/voxmi.h"
#include "../../include/voxmi_batch.h"
#include "vox_internal.h"

using namespace voxmi;

namespace voxmi {

namespace {
std::mutex g_blockMu;
std::multimap<size_t, void*> g_blocks;   // cached device blocks by size
}  // namespace

void* block_take(size_t bytes, size_t* cap) {
    std::lock_guard<std::mutex> lk(g_blockMu);
    auto it = g_blocks.lower_bound(bytes);
    if (it == g_blocks.end() || it->first > bytes + bytes / 4) return nullptr;
    void* p = it->second;
    *cap = it->first;
    g_blocks.erase(it);
    return p;
}

void block_give(void* p, size_t cap) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_blockMu);
    g_blocks.emplace(cap, p);
}

void block_flush() {
    std::lock_guard<std::mutex> lk(g_blockMu);
    for (auto& b : g_blocks) (void)hipFree(b.second);
    g_blocks.clear();
}

}  // namespace voxmi

namespace {

struct HostGroup {
    vox_group desc;
    bool device = false;
    // vox_add_group_gather: desc.ntri rows idx[k] - base of the device arrays
    const uint32_t* idx = nullptr;
    uint32_t base = 0;
};

struct LevelBufs {
    LevelMap map{};
    uint64_t active = 0;
    int aliasOf = -1;     // level whose tiles hold this level's voxels
    DevBuf mark, excl, tileIdx, list, col, qn, nrm;
};

// A level as the kernels see it.  In the reference's mip recursion level
// 2j (j >= 1) is built from level 2j-2 with the same 2x2x2 addressing that
// built level 2j-1 from the same source, so it equals level 2j-1 on its
// smaller extent; such a level reads level 2j-1's tiles with its own dims.
struct LevelView {
    LevelMap map;
    uint64_t active;
    int32_t* list;
    uint32_t* col;
    uint32_t* qn;
    float* nrm;
};

struct DepthBufs {
    uint32_t Rd = 0;
    uint64_t nodes = 0, bricks = 0;
    DevBuf types, cc, bidx, brickNode, bcol, bnrm;
    uint64_t off = 0;   // first node of the depth in the all-depth flag array
    DevBuf ccol, cnrm;  // block-compressed bricks (compressed output)
};

struct TimedLaunch {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct vox_ctx {
    uint32_t N[3] = {0, 0, 0}, B[3] = {0, 0, 0}, chunk[3] = {0, 0, 0};
    uint32_t extra[3] = {0, 0, 0}, S[3] = {0, 0, 0};
    // Chunk edge per axis as processed (S when the requested chunk is at
    // least the voxel space: the same run), chunks per axis, and whether
    // more than one chunk is in play (generateVoxelsAndOctTree's chunk loop,
    // Voxelizer.cu 2285-2309).
    uint32_t C[3] = {0, 0, 0}, nc[3] = {1, 1, 1};
    bool multi = false;
    DevBuf chunkRank;                       // visit rank per chunk
    DevBuf keysA, keysB, nodesA, nodesB, sortTmp;   // brick order with several chunks
    int L = 0;
    bool valid = false;
    bool inited = false;
    bool voxMem = false;
    bool triMem = false;
    bool prepared = false;
    bool haveResult = false;
    bool compressed = false;  // the last generate produced compressed bricks
    // Replay (graph capture) pass: no host round trips; the sizes are the
    // ones the last eager pass measured.
    bool replay = false;
    bool planned = false;
    uint64_t planPairs = 0;
    uint32_t planAct[VOX_MAX_LEVELS] = {};
    uint32_t planNb[VOX_MAX_LEVELS] = {};
    uint32_t options = 0;
    // vox_generate_async: the worker thread owns the handle while running.
    std::thread worker;
    std::atomic<int> async{0};  // 0 none, 1 running, 2 finished (not waited)
    int asyncRet = 0;
    bool workReleased = false;  // vox_release_work_memory freed the level tiles
    float p[3] = {0, 0, 0}, dP[3] = {1, 1, 1}, offP[3] = {0, 0, 0};
    std::stringstream err;
    std::string errCache;
    std::vector<HostGroup> groups;
    hipStream_t st = nullptr;
    // work of a generate that needs none of the work queued beside it on st
    // (occupancy bitmap, collapse), joined back into st before it returns
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;

    // triangle memory
    uint64_t T = 0;
    DevBuf verts, fnorm, bbox, vn, uv, img, grpFill, gbounds;
    std::vector<GroupArgs> gargs;
    // binning
    uint64_t numBins = 0, pairs = 0;
    DevBuf binCount, binStart, binTri, scanTmp;
    // node flags of every depth, depth-major, and their exclusive scan
    DevBuf nodeFlag, nodeExcl;
    uint64_t allNodes = 0;
    // levels
    VoxParams P{};
    std::vector<LevelBufs> lv;
    DevBuf occ0, hit0, sat, work, workCount, rec, recCount;
    DevBuf fill0;   // level-0 voxels reached by a terrain fill, 16 words per active tile
    unsigned persistentBlocks = 2048;
    std::vector<DepthBufs> tree;
    DevBuf bitmap;
    DevBuf counters;  // small host-read counters
    vox_stats stats{};
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> eventPool;
    size_t eventNext = 0;
    hipEvent_t evStart = nullptr, evStop = nullptr;

    bool fail(const std::string& m) {
        err << m << std::endl;
        return false;
    }
    bool check(hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        err << what << " failed: " << hipGetErrorString(e) << std::endl;
        return false;
    }
    hipEvent_t ev() {
        if (eventNext == eventPool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            eventPool.push_back(e);
        }
        return eventPool[eventNext++];
    }
    // Record a named span around launches when timing is on.
    template <class F> void run(const std::string& name, F&& f) {
        if ((options & VOX_OPT_TIMING) && !replay) {
            hipEvent_t a = ev(), b = ev();
            (void)hipEventRecord(a, st);
            f();
            (void)hipEventRecord(b, st);
            timed.push_back({name, a, b});
        } else {
            f();
        }
    }
};

namespace {

// The handle whose asynchronous generate this thread is running.
thread_local vox_ctx* t_worker = nullptr;

// False while a vox_generate_async of h runs on another thread: every entry
// point but vox_wait and vox_destroy then refuses the handle.
bool idle(vox_ctx* h) { return h->async.load(std::memory_order_acquire) != 1 || t_worker == h; }

LevelView view(vox_ctx* h, int k) {
    LevelBufs& l = h->lv[k];
    LevelBufs& b = h->lv[l.aliasOf >= 0 ? l.aliasOf : k];
    LevelView v;
    v.map = b.map;
    for (int a = 0; a < 3; ++a) v.map.dims[a] = l.map.dims[a];
    v.active = b.active;
    v.list = b.list.as<int32_t>();
    v.col = b.col.as<uint32_t>();
    v.qn = b.qn.as<uint32_t>();
    v.nrm = b.nrm.as<float>();
    return v;
}

ChunkGeom chunkGeom(const vox_ctx* h) {
    ChunkGeom g{};
    for (int a = 0; a < 3; ++a) {
        g.C[a] = h->C[a];
        g.nc[a] = h->nc[a];
    }
    g.rank = h->chunkRank.as<uint32_t>();
    return g;
}

inline float gclampf(float x, float lo, float hi) {
    float m = x < hi ? x : hi;
    return m > lo ? m : lo;
}
inline uint64_t trunc_u(float f) {
    if (!(f == f) || f <= 0.0f) return 0;
    if (f >= 1.8e19f) return UINT64_MAX;
    return (uint64_t)f;
}

// Word of h->counters where the scans of a generate report a wait that
// timed out (zeroed by the setup pass, read at the generate's last sync).
constexpr uint32_t kScanErrWord = 128;

bool scan(vox_ctx* h, const uint32_t* in, uint32_t* out, uint64_t n) {
    if (!ensure_scan_tmp(h->scanTmp, n, h->st, h->err, "scan scratch")) return false;
    h->stats.scan_values += n;
    launch_scan(h->st, in, out, n, h->scanTmp.as<uint32_t>(),
                h->counters.as<uint32_t>() + kScanErrWord);
    return true;
}

// Voxels of the x axis examined by ComputeOctTreeNodeType for one level: the
// launch is split in two along x when the extent exceeds 128 voxels and the
// halved block count is only rounded up when odd (Voxelizer.cu 1037-1120),
// which can leave the last x block uncovered.
uint32_t type_x_limit(uint32_t Ck, uint32_t border, uint32_t fullVoxDim, const uint32_t Ckyz[2],
                      const uint32_t borderyz[2], const uint32_t fullyz[2]) {
    uint32_t tpb[3];
    const uint32_t C3[3] = {Ck, Ckyz[0], Ckyz[1]};
    const uint32_t b3[3] = {border, borderyz[0], borderyz[1]};
    const uint32_t f3[3] = {fullVoxDim, fullyz[0], fullyz[1]};
    for (int a = 0; a < 3; ++a) {
        tpb[a] = C3[a];
        uint32_t fvo = tpb[a] + 0u - b3[a];
        if (fvo > f3[a]) tpb[a] = f3[a] - (0u - b3[a]) + 2u;
    }
    uint32_t th = tpb[0], nb = 1;
    if (tpb[0] > 8 || tpb[1] > 8 || tpb[2] > 8) {
        th = std::min(tpb[0], 8u);
        nb = (uint32_t)std::ceil((float)tpb[0] / (float)th);
    }
    uint32_t nexec = 1;
    if (tpb[0] > 128u) {
        nb >>= 1;
        if ((nb & 1u) != 0u) nb += 1;
        nexec = 2;
    }
    return std::min(nexec * nb * th, Ck);
}

void collect_timings(vox_ctx* h) {
    vox_stats& s = h->stats;
    s.nkernels = 0;
    std::map<std::string, int> slot;
    for (auto& t : h->timed) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        auto it = slot.find(t.name);
        int i;
        if (it == slot.end()) {
            if (s.nkernels >= 64) continue;
            i = s.nkernels++;
            slot[t.name] = i;
            std::snprintf(s.kernel_name[i], sizeof(s.kernel_name[i]), "%s", t.name.c_str());
            s.kernel_ms[i] = 0.0;
            s.kernel_launches[i] = 0;
        } else {
            i = it->second;
        }
        s.kernel_ms[i] += ms;
        s.kernel_launches[i] += 1;
    }
    h->timed.clear();
    h->eventNext = 0;
}

}  // namespace

extern "C" {

vox_ctx* vox_create(const uint32_t vox_dim[3], const uint32_t brick_dim[3],
                    const uint32_t chunk_dim[3]) {
    vox_ctx* h = new (std::nothrow) vox_ctx();
    if (!h) return nullptr;
    int L = 0;
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        h->N[a] = vox_dim[a];
        h->B[a] = brick_dim[a];
        if (brick_dim[a] < 2 || vox_dim[a] % brick_dim[a] != 0) {
            h->err << "voxel dimension must be a multiple of a brick dimension >= 2" << std::endl;
            ok = false;
            break;
        }
        uint32_t r = vox_dim[a] / brick_dim[a];
        if (r < 2 || (r & (r - 1)) != 0) {
            // The reference loops forever here (Voxelizer.cu 146-152).
            h->err << "voxel / brick ratio must be a power of two >= 2" << std::endl;
            ok = false;
            break;
        }
        int l = 1;
        while ((brick_dim[a] << (l - 1)) != vox_dim[a]) ++l;
        if (a > 0 && l != L) {
            h->err << "voxel / brick ratio must be equal on all axes" << std::endl;
            ok = false;
            break;
        }
        L = l;
        h->extra[a] = r >> 1;
        h->S[a] = vox_dim[a] + 4 * h->extra[a];
        h->chunk[a] = chunk_dim ? chunk_dim[a] : h->S[a];
        if (h->S[a] > 65535u) {
            h->err << "voxel space above 65535 per axis is not supported" << std::endl;
            ok = false;
            break;
        }
        if (L > VOX_MAX_LEVELS) {
            h->err << "too many octree levels" << std::endl;
            ok = false;
            break;
        }
    }
    // Chunks smaller than the voxel space: the chunk must be a multiple of
    // 2^(levels-1), so that every level's chunk origin (x >> k, Voxelizer.cu
    // 1273-1276) is a whole number of that level's chunk edges, and at most
    // 64 chunks per axis.
    for (int a = 0; a < 3 && ok; ++a) {
        if (h->chunk[a] >= h->S[a]) {
            h->C[a] = h->S[a];
            h->nc[a] = 1;
            continue;
        }
        if (h->chunk[a] == 0 || h->chunk[a] % (1u << (L - 1)) != 0) {
            h->err << "chunk dimension " << h->chunk[a] << " must be a multiple of 2^(levels-1) = "
                   << (1u << (L - 1)) << " or at least the voxel space " << h->S[a]
                   << " (N + 4*((N/B)>>1))" << std::endl;
            ok = false;
            break;
        }
        h->C[a] = h->chunk[a];
        // chunk origins are the multiples of the chunk below N (Voxelizer.cu 2300-2302)
        h->nc[a] = (h->N[a] + h->C[a] - 1) / h->C[a];
        if (h->nc[a] > (uint32_t)kMaxChunksAxis) {
            h->err << "more than " << kMaxChunksAxis << " chunks per axis" << std::endl;
            ok = false;
            break;
        }
        h->multi = true;
    }
    if (ok && (uint64_t)(h->B[0] + 2) * (h->B[1] + 2) * (h->B[2] + 2) >= (1ull << 22)) {
        // brick voxels are indexed with 22-bit quotients in the gather
        h->err << "bricks of more than 2^22 stored voxels ((B+2)^3) are not supported" << std::endl;
        ok = false;
    }
    h->L = L;
    h->valid = ok;
    return h;
}

void vox_deallocate_triangle_memory(vox_ctx* h);
void vox_deallocate_voxel_memory(vox_ctx* h);

void vox_destroy(vox_ctx* h) {
    if (!h) return;
    if (h->worker.joinable()) h->worker.join();
    h->async.store(0);
    vox_deallocate_triangle_memory(h);
    vox_deallocate_voxel_memory(h);
    for (auto e : h->eventPool) (void)hipEventDestroy(e);
    if (h->evStart) (void)hipEventDestroy(h->evStart);
    if (h->evStop) (void)hipEventDestroy(h->evStop);
    if (h->evFork) (void)hipEventDestroy(h->evFork);
    if (h->evJoin) (void)hipEventDestroy(h->evJoin);
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

int vox_set_params(vox_ctx* h, const float p[3], const float delta[3]) {
    if (!h || !h->valid || !idle(h)) return 0;
    for (int a = 0; a < 3; ++a) {
        h->p[a] = p[a];
        h->dP[a] = delta[a];
        // offsetP = p - extra * delta (Voxelizer.cu 160)
        h->offP[a] = p[a] - ((float)h->extra[a] * delta[a]);
    }
    return 1;
}

int vox_set_options(vox_ctx* h, uint32_t options) {
    if (!h || !idle(h)) return 0;
    h->options = options;
    return 1;
}

static int add_group(vox_ctx* h, const vox_group* g, bool device) {
    if (!h || !h->valid || !g || !idle(h)) return 0;
    if (!g->verts && g->ntri > 0) return h->fail("group has no vertices"), 0;
    if (g->rgba8 && (g->width <= 0 || g->height <= 0)) return h->fail("group image has no size"), 0;
    if (h->T + g->ntri > 0x7fffffffull) return h->fail("more than 2^31 triangles"), 0;
    HostGroup hg;
    hg.desc = *g;
    hg.device = device;
    h->groups.push_back(hg);
    h->T += g->ntri;
    h->triMem = false;
    h->prepared = false;
    return 1;
}
int vox_add_group(vox_ctx* h, const vox_group* g) { return add_group(h, g, false); }
int vox_add_group_device(vox_ctx* h, const vox_group* g) { return add_group(h, g, true); }

int vox_add_group_gather(vox_ctx* h, const vox_group* g, const uint32_t* idx, uint64_t n,
                         uint32_t base) {
    if (!h || !h->valid || !g || !idle(h)) return 0;
    if (n && !idx) return h->fail("gathered group has no index list"), 0;
    if (n > g->ntri) return h->fail("gathered group selects more triangles than it has"), 0;
    vox_group d = *g;
    d.ntri = n;
    if (!add_group(h, &d, true)) return 0;
    h->groups.back().idx = idx;
    h->groups.back().base = base;
    return 1;
}

int vox_init(vox_ctx* h) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (h->inited) return 1;
    if (!h->check(hipSetDevice(0), "hipSetDevice")) return 0;
    if (!h->check(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking), "hipStreamCreate") ||
        !h->check(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking), "hipStreamCreate") ||
        !h->check(hipEventCreateWithFlags(&h->evFork, hipEventDisableTiming), "hipEventCreate") ||
        !h->check(hipEventCreateWithFlags(&h->evJoin, hipEventDisableTiming), "hipEventCreate"))
        return 0;
    (void)hipEventCreate(&h->evStart);
    (void)hipEventCreate(&h->evStop);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess && cus > 0)
        h->persistentBlocks = (unsigned)cus * 8u;
    h->inited = true;
    return 1;
}

int vox_allocate_voxel_memory(vox_ctx* h) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (!h->inited && !vox_init(h)) return 0;
    // Tile maps and per-depth node arrays are sized here; tile pools grow on
    // demand in vox_generate.
    VoxParams& P = h->P;
    for (int a = 0; a < 3; ++a) {
        P.S[a] = h->S[a];
        P.TG[a] = (h->S[a] + kTile - 1) / kTile;
        P.C[a] = h->C[a];
    }
    P.numTiles = (uint64_t)P.TG[0] * P.TG[1] * P.TG[2];
    h->lv.assign(h->L, LevelBufs());
    for (int k = 0; k < h->L; ++k) {
        LevelBufs& l = h->lv[k];
        for (int a = 0; a < 3; ++a) {
            // level 0 is the voxel space; level k >= 1 holds each chunk's
            // own mip chain side by side (S >> k with one chunk)
            l.map.dims[a] = k == 0 ? h->S[a] : h->nc[a] * (h->C[a] >> k);
            l.map.TG[a] = (l.map.dims[a] + kTile - 1) / kTile;
        }
        l.map.numTiles = (uint64_t)l.map.TG[0] * l.map.TG[1] * l.map.TG[2];
        uint64_t nt = l.map.numTiles;
        if (!l.mark.ensure(nt * 4, h->err, "tile marks") ||
            !l.excl.ensure((nt + 1) * 4, h->err, "tile scan") ||
            !l.tileIdx.ensure(nt * 4, h->err, "tile map") ||
            !l.list.ensure(nt * 4, h->err, "tile list"))
            return 0;
        l.map.tileIdx = l.tileIdx.as<int32_t>();
    }
    h->tree.assign(h->L, DepthBufs());
    h->allNodes = 0;
    for (int d = 0; d < h->L; ++d) {
        DepthBufs& t = h->tree[d];
        t.Rd = 1u << d;
        t.nodes = (uint64_t)t.Rd * t.Rd * t.Rd;
        t.off = h->allNodes;
        h->allNodes += t.nodes;
        if (!t.types.ensure(t.nodes, h->err, "node types") ||
            !t.cc.ensure(t.nodes * 4, h->err, "node colours") ||
            !t.bidx.ensure(t.nodes * 4, h->err, "brick index"))
            return 0;
    }
    if (!h->nodeFlag.ensure(h->allNodes * 4, h->err, "node flags") ||
        !h->nodeExcl.ensure((h->allNodes + 1) * 4, h->err, "node scan"))
        return 0;
    if (!h->counters.ensure(4096, h->err, "counters")) return 0;
    if (!h->sat.ensure(8, h->err, "sat counter")) return 0;
    if (h->multi) {
        // Chunk visit order: breadth first from the origin, each chunk
        // queueing its +x/+y/+z neighbours in x, y, z loop order
        // (Voxelizer.cu 2285-2309); rank per chunk (i, j, l).
        const uint32_t* nc = h->nc;
        std::vector<uint32_t> rank((size_t)nc[0] * nc[1] * nc[2], 0xffffffffu);
        std::vector<std::array<uint32_t, 3>> q;
        q.push_back({0, 0, 0});
        rank[0] = 0;
        for (size_t head = 0; head < q.size(); ++head) {
            const auto c = q[head];
            for (uint32_t dx = 0; dx < 2; ++dx)
                for (uint32_t dy = 0; dy < 2; ++dy)
                    for (uint32_t dz = 0; dz < 2; ++dz) {
                        const uint32_t n[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                        if (n[0] >= nc[0] || n[1] >= nc[1] || n[2] >= nc[2]) continue;
                        const size_t i = ((size_t)n[2] * nc[1] + n[1]) * nc[0] + n[0];
                        if (rank[i] != 0xffffffffu) continue;
                        rank[i] = (uint32_t)q.size();
                        q.push_back({n[0], n[1], n[2]});
                    }
        }
        uint64_t maxNodes = 0;
        for (int d = 0; d < h->L; ++d) maxNodes = std::max(maxNodes, h->tree[d].nodes);
        if (!h->chunkRank.ensure(rank.size() * 4, h->err, "chunk ranks") ||
            !h->check(hipMemcpy(h->chunkRank.p, rank.data(), rank.size() * 4, hipMemcpyHostToDevice),
                      "chunk ranks") ||
            !h->keysA.ensure(h->allNodes * 8, h->err, "brick keys") ||
            !h->keysB.ensure(maxNodes * 8, h->err, "brick keys") ||
            !h->nodesA.ensure(h->allNodes * 4, h->err, "brick nodes") ||
            !h->nodesB.ensure(maxNodes * 4, h->err, "brick nodes"))
            return 0;
        size_t tmpBytes = 0;
        if (!h->check(chunk_sort_pairs(h->st, nullptr, &tmpBytes, nullptr, nullptr, nullptr, nullptr,
                                       (uint32_t)maxNodes, 64),
                      "sort size") ||
            !h->sortTmp.ensure(tmpBytes, h->err, "sort scratch"))
            return 0;
    }
    h->voxMem = true;
    return 1;
}

void vox_deallocate_voxel_memory(vox_ctx* h) {
    if (!h || !idle(h)) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (auto& l : h->lv) {
        l.mark.release(); l.excl.release(); l.tileIdx.release(); l.list.release();
        l.col.release(); l.qn.release(); l.nrm.release();
    }
    h->lv.clear();
    for (auto& t : h->tree) {
        t.types.release(); t.cc.release(); t.bidx.release();
        t.brickNode.release(); t.bcol.release(); t.bnrm.release();
    }
    h->tree.clear();
    h->nodeFlag.release(); h->nodeExcl.release();
    h->occ0.release(); h->hit0.release(); h->fill0.release(); h->sat.release();
    h->bitmap.release();
    h->work.release(); h->workCount.release(); h->rec.release(); h->recCount.release();
    h->counters.release(); h->binCount.release(); h->binStart.release(); h->binTri.release();
    h->scanTmp.release();
    h->chunkRank.release(); h->keysA.release(); h->keysB.release(); h->nodesA.release();
    h->nodesB.release(); h->sortTmp.release();
    h->voxMem = false;
    h->haveResult = false;
}

int vox_allocate_triangle_memory(vox_ctx* h) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (!h->inited && !vox_init(h)) return 0;
    uint64_t T = h->T, nvn = 0, nuv = 0, nimg = 0;
    for (auto& g : h->groups) {
        if (g.desc.vnormals) nvn += g.desc.ntri;
        if (g.desc.uvs && g.desc.rgba8) {
            nuv += g.desc.ntri;
            nimg += (uint64_t)g.desc.width * g.desc.height * 4;
        }
    }
    if (!h->gbounds.ensure(h->groups.size() * 24 + 24, h->err, "group bounds") ||
        !h->verts.ensure(T * 36, h->err, "vertex buffer") ||
        !h->fnorm.ensure(T * 12, h->err, "face normals") ||
        !h->bbox.ensure(T * 12, h->err, "bounds") ||
        !h->vn.ensure(nvn * 36, h->err, "vertex normals") ||
        !h->uv.ensure(nuv * 24, h->err, "uvs") ||
        !h->img.ensure(nimg, h->err, "images") ||
        !h->grpFill.ensure(h->groups.size() * 4 + 4, h->err, "group flags"))
        return 0;
    // Concatenate groups (allocateTriangleMemory, Voxelizer.cu 346-457).
    uint64_t to = 0, vo = 0, uo = 0, io = 0;
    h->gargs.clear();
    for (auto& g : h->groups) {
        const vox_group& d = g.desc;
        hipMemcpyKind kind = g.device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        GroupArgs ga{};
        ga.triOff = (uint32_t)to;
        ga.ntri = (uint32_t)d.ntri;
        const bool gather = g.idx != nullptr;
        if (gather) {
            // rows picked by index, and the group's bounds from them (the
            // tiler's per-tile triBBox, main.cu 3868-3870)
            int32_t* gb = h->gbounds.as<int32_t>() + 6 * (&g - h->groups.data());
            launch_fill_u32(h->st, (uint32_t*)gb, 3, 0x7fffffffu);
            launch_fill_u32(h->st, (uint32_t*)gb + 3, 3, 0x80000000u);
            launch_gather_rows(h->st, d.verts, g.idx, d.ntri, g.base, 9, h->verts.as<float>() + 9 * to);
            launch_tri_bounds(h->st, h->verts.as<float>() + 9 * to, d.ntri, gb);
        } else if (d.ntri && !h->check(hipMemcpyAsync(h->verts.as<float>() + 9 * to, d.verts,
                                                      d.ntri * 36, kind, h->st), "vertex copy")) {
            return 0;
        }
        if (d.vnormals) {
            if (gather)
                launch_gather_rows(h->st, d.vnormals, g.idx, d.ntri, g.base, 9,
                                   h->vn.as<float>() + 9 * vo);
            else if (d.ntri && !h->check(hipMemcpyAsync(h->vn.as<float>() + 9 * vo, d.vnormals,
                                                        d.ntri * 36, kind, h->st), "normal copy"))
                return 0;
            ga.vn = h->vn.as<float>() + 9 * vo;
            ga.hasVN = 1;
            vo += d.ntri;
        }
        if (d.uvs && d.rgba8) {
            if (gather)
                launch_gather_rows(h->st, d.uvs, g.idx, d.ntri, g.base, 6, h->uv.as<float>() + 6 * uo);
            else if (d.ntri && !h->check(hipMemcpyAsync(h->uv.as<float>() + 6 * uo, d.uvs,
                                                        d.ntri * 24, kind, h->st), "uv copy"))
                return 0;
            uint64_t ib = (uint64_t)d.width * d.height * 4;
            if (!h->check(hipMemcpyAsync(h->img.as<uint8_t>() + io, d.rgba8, ib, kind, h->st),
                          "image copy"))
                return 0;
            ga.uv = h->uv.as<float>() + 6 * uo;
            ga.img = h->img.as<uint8_t>() + io;
            ga.w = d.width;
            ga.h = d.height;
            ga.addr0 = d.addr0;
            ga.addr1 = d.addr1;
            ga.textured = 1;
            uo += d.ntri;
            io += ib;
        }
        ga.terrain = d.is_terrain ? 1 : 0;
        h->gargs.push_back(ga);
        to += d.ntri;
    }
    if (!h->check(hipGetLastError(), "triangle gather") ||
        !h->check(hipStreamSynchronize(h->st), "triangle upload"))
        return 0;
    bool anyGather = false;
    for (auto& g : h->groups) anyGather |= g.idx != nullptr;
    if (anyGather) {
        std::vector<int32_t> gb(6 * h->groups.size());
        if (!h->check(hipMemcpy(gb.data(), h->gbounds.p, 4 * gb.size(), hipMemcpyDeviceToHost),
                      "group bounds"))
            return 0;
        for (size_t k = 0; k < h->groups.size(); ++k) {
            HostGroup& g = h->groups[k];
            if (!g.idx || g.desc.ntri == 0) continue;
            for (int a = 0; a < 6; ++a) {
                const int32_t o = gb[6 * k + a];
                const int32_t bits = o >= 0 ? o : (int32_t)(0x80000000u - (uint32_t)(o + 1));
                float f;
                std::memcpy(&f, &bits, 4);
                (a < 3 ? g.desc.bbox_min[a] : g.desc.bbox_max[a - 3]) = f;
            }
        }
    }
    h->triMem = true;
    h->prepared = false;
    return 1;
}

void vox_deallocate_triangle_memory(vox_ctx* h) {
    if (!h || !idle(h)) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    h->verts.release(); h->fnorm.release(); h->bbox.release(); h->vn.release(); h->gbounds.release();
    h->uv.release(); h->img.release(); h->grpFill.release();
    h->groups.clear();
    h->gargs.clear();
    h->T = 0;
    h->triMem = false;
    h->prepared = false;
}

int vox_compute_edges_face_normals_and_bounds(vox_ctx* h) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (!h->voxMem && !vox_allocate_voxel_memory(h)) return 0;
    if (!h->triMem && !vox_allocate_triangle_memory(h)) return 0;
    VoxParams& P = h->P;
    for (int a = 0; a < 3; ++a) {
        P.offP[a] = h->offP[a];
        P.dP[a] = h->dP[a];
    }
    const int G = (int)h->groups.size();
    // Group voxel bounds (addTriangleGroup, Voxelizer.h 137-144) and chunk
    // overlap (Voxelizer.cu 2341-2353).
    std::vector<int32_t> fill(G + 1, 0);
    for (int g = 0; g < G; ++g) {
        const vox_group& d = h->groups[g].desc;
        GroupArgs& ga = h->gargs[g];
        uint32_t mn[3], mx[3];
        for (int a = 0; a < 3; ++a) {
            float lo = gclampf((d.bbox_min[a] - h->offP[a]) / h->dP[a], 0.0f, (float)h->S[a]);
            float hi = gclampf((d.bbox_max[a] - h->offP[a]) / h->dP[a], 0.0f, (float)h->S[a]);
            mn[a] = (uint32_t)trunc_u(lo);
            mx[a] = (uint32_t)trunc_u(hi);
        }
        // The chunks the bounds overlap (a chunk [iC, iC + C - 1] with i <
        // nc; one chunk of S voxels when it holds the space): x and y reach
        // those chunks, z the group's own range inside them.
        uint32_t lo[3], hi[3];
        bool any = true;
        for (int a = 0; a < 3; ++a) {
            const uint32_t cov = h->nc[a] * h->C[a];
            if (mn[a] > cov - 1) any = false;
            const uint32_t i0 = mn[a] / h->C[a], i1 = std::min(mx[a] / h->C[a], h->nc[a] - 1);
            lo[a] = i0 * h->C[a];
            hi[a] = std::min(std::min((i1 + 1) * h->C[a], cov), h->S[a]) - 1;
        }
        ga.overlaps = any ? 1 : 0;
        ga.xStart = lo[0];
        ga.xEnd = hi[0];
        ga.yStart = lo[1];
        ga.yEnd = hi[1];
        ga.zStart = mn[2];
        ga.zEnd = std::min(mx[2], hi[2]);
        fill[g] = (ga.overlaps && ga.textured && ga.terrain) ? 1 : 0;
    }
    // (replay: the flags are already on the device from the eager pass)
    if (!h->replay &&
        !h->check(hipMemcpyAsync(h->grpFill.p, fill.data(), 4 * (G + 1), hipMemcpyHostToDevice, h->st),
                  "group flags"))
        return 0;
    h->numBins = (uint64_t)std::max(G, 1) * P.numTiles;
    if (!h->binCount.ensure((h->numBins + 1) * 4, h->err, "bin counts") ||
        !h->binStart.ensure((h->numBins + 1) * 4, h->err, "bin offsets"))
        return 0;
    uint32_t* cnt = h->binCount.as<uint32_t>();
    uint32_t* start = h->binStart.as<uint32_t>();
    h->run("fill", [&] {
        const FillSeg z[2] = {{cnt, h->numBins, 0u},
                              {h->counters.as<uint32_t>() + kScanErrWord, 1, 0u}};
        launch_fill_segs(h->st, z, 2);
    });
    h->run("setup", [&] {
        launch_setup(h->st, P, h->gargs, h->verts.as<float>(), h->fnorm.as<float>(),
                     h->bbox.as<uint32_t>(), cnt);
    });
    // the scan leaves the counts zeroed for the scatter pass to count again
    if (!ensure_scan_tmp(h->scanTmp, h->numBins, h->st, h->err, "scan scratch")) return 0;
    h->run("scan", [&] {
        launch_scan_clear(h->st, cnt, start, h->numBins, h->scanTmp.as<uint32_t>(),
                          h->counters.as<uint32_t>() + kScanErrWord);
    });
    uint32_t total = 0;
    if (h->replay) {
        total = (uint32_t)h->planPairs;
    } else {
        if (!h->check(hipMemcpyAsync(&total, start + h->numBins, 4, hipMemcpyDeviceToHost, h->st),
                      "bin total") ||
            !h->check(hipStreamSynchronize(h->st), "setup"))
            return 0;
        h->planPairs = total;
    }
    h->pairs = total;
    if (!h->binTri.ensure((uint64_t)total * 4, h->err, "bin entries")) return 0;
    h->run("scatter", [&] {
        launch_scatter(h->st, P, h->gargs, h->bbox.as<uint32_t>(), start, cnt,
                       h->binTri.as<uint32_t>());
    });
    h->prepared = true;
    return 1;
}

// the file writers (gvformat.cpp)
int vox_write_files_impl(vox_ctx* h, const char* out_dir, const float p[3], const float delta[3],
                         const uint32_t N[3], int binary, const uint32_t extra[3],
                         const uint32_t chunk[3], int multi);

namespace {

// vox_generate in four parts, so a batch capture can put one persistent
// voxelize launch of all its jobs between every job's tile maps and its
// colour pass: the tile maps and buffers (gen_front), per overlapping group
// its work list (gen_group_tiles), voxelize (gen_voxelize, or its arguments
// from gen_voxelize_job) and colour + fill (gen_group_tail), then the mips,
// octree and bricks (gen_back).
int first_overlap(const vox_ctx* h) {
    for (size_t g = 0; g < h->gargs.size(); ++g)
        if (h->gargs[g].overlaps) return (int)g;
    return -1;
}

int gen_front(vox_ctx* h) {
    const int G = (int)h->groups.size();
    const int L = h->L;
    VoxParams& P = h->P;
    hipStream_t st = h->st;
    if (!h->replay) {
        std::memset(&h->stats, 0, sizeof(h->stats));
        (void)hipEventRecord(h->evStart, st);
    }

    // Active tiles: level 0 from the bins, level k from level src(k).
    LevelBufs& l0 = h->lv[0];
    h->run("activity", [&] {
        launch_activity0(st, P, G, h->grpFill.as<int32_t>(), h->binStart.as<uint32_t>(),
                         l0.mark.as<uint32_t>());
    });
    const bool chain = (h->options & VOX_OPT_MIP_CHAIN) != 0;
    auto srcOf = [&](int k) { return chain ? k - 1 : (k == 1 ? 0 : k - 2); };
    uint32_t* ctr = h->counters.as<uint32_t>();
    // Each level is compacted before the next one marks its parents; the
    // dense tile list is sized for every tile of the level.
    for (int k = 0; k < L; ++k) {
        LevelBufs& l = h->lv[k];
        // (each chunk has its own chain: level 2j is not level 2j - 1 then)
        l.aliasOf = (!chain && !h->multi && k >= 2 && (k % 2) == 0) ? k - 1 : -1;
        if (l.aliasOf >= 0) continue;
        if (k > 0) {
            LevelView s = view(h, srcOf(k));
            MipShift shift;
            if (h->multi && mip_shift_for(s.map, l.map, chunkGeom(h), srcOf(k), k, shift))
                h->run("mark", [&] {
                    launch_mark_children_shifted(st, s.map, l.map, shift, l.mark.as<uint32_t>());
                });
            else if (h->multi)
                h->run("mark", [&] {
                    launch_chunk_mark(st, s.map, l.map, chunkGeom(h), srcOf(k), k,
                                      l.mark.as<uint32_t>());
                });
            else
                h->run("mark", [&] { launch_mark_children(st, s.map, l.map, l.mark.as<uint32_t>()); });
        }
        h->run("scan", [&] { scan(h, l.mark.as<uint32_t>(), l.excl.as<uint32_t>(), l.map.numTiles); });
        h->run("compact", [&] {
            launch_compact(st, l.map.numTiles, l.mark.as<uint32_t>(), l.excl.as<uint32_t>(),
                           l.tileIdx.as<int32_t>(), l.list.as<int32_t>(), ctr + k);
        });
    }
    uint32_t act[VOX_MAX_LEVELS] = {0};
    if (h->replay) {
        std::memcpy(act, h->planAct, sizeof(act));
    } else {
        if (!h->check(hipMemcpyAsync(act, ctr, 4 * L, hipMemcpyDeviceToHost, st), "tile counts") ||
            !h->check(hipStreamSynchronize(st), "tile maps"))
            return 0;
        std::memcpy(h->planAct, act, sizeof(act));
    }
    for (int k = 0; k < L; ++k) {
        LevelBufs& l = h->lv[k];
        if (l.aliasOf >= 0) {
            l.active = 0;
            h->stats.active_tiles[k] = h->lv[l.aliasOf].active;
            continue;
        }
        l.active = act[k];
        h->stats.active_tiles[k] = act[k];
        if (!l.col.ensure((uint64_t)act[k] * kTileVox * 4, h->err, "tile colours") ||
            !l.qn.ensure((uint64_t)act[k] * kTileVox * 4, h->err, "tile brick normals") ||
            !l.nrm.ensure((uint64_t)act[k] * kTileVox * 12, h->err, "tile normals"))
            return 0;
    }
    if (kRecSegs * rec_seg_cap(l0.active, vox_grid(l0.active, h->persistentBlocks)) >=
        (1ull << 32))
        return h->fail("more hit records than 32-bit indices address; use fewer persistent "
                       "workgroups or a smaller grid"), 0;
    if (!h->occ0.ensure((uint64_t)l0.active * kTileWords * 4, h->err, "occupancy words") ||
        !h->fill0.ensure((uint64_t)l0.active * kTileWords * 4, h->err, "filled-voxel words") ||
        !h->hit0.ensure((uint64_t)l0.active * kTileWords * 4, h->err, "hit words") ||
        !h->work.ensure((uint64_t)l0.active * 16, h->err, "tile work list") ||
        !h->workCount.ensure(4, h->err, "tile work count") ||
        !h->rec.ensure(kRecSegs * rec_seg_cap(l0.active, vox_grid(l0.active, h->persistentBlocks)) *
                           32,
                       h->err, "hit records") ||
        !h->recCount.ensure(4 * kRecSegs, h->err, "hit record counts"))
        return 0;
    unsigned long long* satc = h->sat.as<unsigned long long>();
    h->run("fill", [&] {
        const FillSeg z[3] = {{h->occ0.as<uint32_t>(), l0.active * kTileWords, 0u},
                              {h->fill0.as<uint32_t>(), l0.active * kTileWords, 0u},
                              {reinterpret_cast<uint32_t*>(satc), 2, 0u}};
        launch_fill_segs(st, z, 3);
    });
    // Level 0 stores only hit or filled voxels: the others read as zero
    // wherever level 0 is read (LevelMap::defOcc / defFill), so no tile is
    // cleared first and the terrain fill writes no zeros.
    l0.map.defOcc = h->occ0.as<uint32_t>();
    l0.map.defFill = h->fill0.as<uint32_t>();

    return 1;
}

void gen_group_tiles(vox_ctx* h, int g) {
    LevelBufs& l0 = h->lv[0];
    h->run("group_tiles", [&] {
        launch_group_tiles(h->st, h->P, g, l0.active, l0.list.as<int32_t>(),
                           h->binStart.as<uint32_t>(), h->work.as<uint4>(),
                           h->workCount.as<uint32_t>(), h->recCount.as<uint32_t>());
    });
}

void gen_voxelize(vox_ctx* h, int g) {
    LevelBufs& l0 = h->lv[0];
    h->run("voxelize_g" + std::to_string(g), [&] {
        launch_voxelize(h->st, h->P, h->gargs[g], g, g == first_overlap(h) ? 1 : 0, l0.active,
                        l0.list.as<int32_t>(), h->verts.as<float>(), h->fnorm.as<float>(),
                        h->bbox.as<uint32_t>(), h->binStart.as<uint32_t>(),
                        h->binTri.as<uint32_t>(), h->occ0.as<uint32_t>(), h->hit0.as<uint32_t>(),
                        h->sat.as<unsigned long long>(), h->work.as<uint4>(),
                        h->workCount.as<uint32_t>(), h->rec.as<uint4>(),
                        h->recCount.as<uint32_t>(), h->persistentBlocks);
    });
}

VoxUnitJob gen_voxelize_job(vox_ctx* h, int g) {
    LevelBufs& l0 = h->lv[0];
    return voxelize_job(h->P, h->gargs[g], g == first_overlap(h) ? 1 : 0, l0.active,
                        h->verts.as<float>(), h->fnorm.as<float>(), h->bbox.as<uint32_t>(),
                        h->binTri.as<uint32_t>(), h->occ0.as<uint32_t>(), h->hit0.as<uint32_t>(),
                        h->work.as<uint4>(), h->workCount.as<uint32_t>(), h->rec.as<uint4>(),
                        h->recCount.as<uint32_t>(), h->persistentBlocks);
}

void gen_group_tail(vox_ctx* h, int g) {
    LevelBufs& l0 = h->lv[0];
    const GroupArgs& ga = h->gargs[g];
    const int init = g == first_overlap(h) ? 1 : 0;
    hipStream_t st = h->st;
    const VoxParams& P = h->P;
    h->run("colour_g" + std::to_string(g), [&] {
        launch_colour(st, P, ga, init, l0.active, h->rec.as<uint4>(), h->recCount.as<uint32_t>(),
                      l0.list.as<int32_t>(), h->verts.as<float>(), h->fnorm.as<float>(),
                      l0.col.as<uint32_t>(), l0.qn.as<uint32_t>(), l0.nrm.as<float>(),
                      h->fill0.as<uint32_t>(), h->persistentBlocks);
    });
    if (ga.textured && ga.terrain) {
        h->run("fill_terrain", [&] {
            launch_fill(st, P, g, init, l0.active, l0.list.as<int32_t>(),
                        l0.tileIdx.as<int32_t>(), h->binStart.as<uint32_t>(),
                        l0.col.as<uint32_t>(), l0.qn.as<uint32_t>(), l0.nrm.as<float>(),
                        h->hit0.as<uint32_t>(), h->fill0.as<uint32_t>());
        });
    }
}

int gen_back(vox_ctx* h, const char* out_dir, int binary) {
    const int L = h->L;
    VoxParams& P = h->P;
    hipStream_t st = h->st;
    LevelBufs& l0 = h->lv[0];
    const bool chain = (h->options & VOX_OPT_MIP_CHAIN) != 0;
    auto srcOf = [&](int k) { return chain ? k - 1 : (k == 1 ? 0 : k - 2); };
    uint32_t* ctr = h->counters.as<uint32_t>();
    unsigned long long* satc = h->sat.as<unsigned long long>();
    const int firstOverlap = first_overlap(h);
    // The occupancy bitmap needs level 0 only, and the collapse the node
    // types only: both go to the side stream, where they fill the wait for
    // the brick counts and run beside the brick gather, and are joined back
    // into st before the counters are read (on every return).  A timing
    // run, whose spans are events on st, and a captured generate keep
    // everything on st.
    const bool overlap = h->side && !h->replay && !(h->options & VOX_OPT_TIMING);
    hipStream_t aux = overlap ? h->side : st;
    struct SideJoin {
        vox_ctx* h;
        bool forked = false;
        void fork() {
            (void)hipEventRecord(h->evFork, h->st);
            (void)hipStreamWaitEvent(h->side, h->evFork, 0);
            forked = true;
        }
        void join() {
            if (!forked) return;
            (void)hipEventRecord(h->evJoin, h->side);
            (void)hipStreamWaitEvent(h->st, h->evJoin, 0);
            forked = false;
        }
        ~SideJoin() { join(); }
    } side{h};
    // Dense occupancy bitmap (launched once the node types are queued).
    const uint64_t nvox = (uint64_t)h->S[0] * h->S[1] * h->S[2];
    const uint64_t words = (nvox + 31) / 32;
    const bool wantOcc = !(h->options & VOX_OPT_NO_OCCUPANCY);
    if (wantOcc && !h->bitmap.ensure(words * 4, h->err, "occupancy bitmap")) return 0;
    auto occupancy = [&] {
        if (!wantOcc) return;
        if (overlap) side.fork();
        bool rows = false;
        h->run("occupancy", [&] {
            rows = launch_occupancy_rows(aux, P, l0.tileIdx.as<int32_t>(), h->occ0.as<uint32_t>(),
                                         h->bitmap.as<uint32_t>());
        });
        if (!rows) {
            h->run("fill", [&] { launch_fill_u32(aux, h->bitmap.as<uint32_t>(), words, 0u); });
            h->run("occupancy", [&] {
                launch_occupancy(aux, P, l0.active, l0.list.as<int32_t>(), h->occ0.as<uint32_t>(),
                                 h->bitmap.as<uint32_t>());
            });
        }
    };
    // Mip chain (computeMipMaps, Voxelizer.cu 491-594).
    for (int k = 1; k < L; ++k) {
        if (h->lv[k].aliasOf >= 0) continue;
        LevelView s = view(h, srcOf(k));
        LevelView l = view(h, k);
        MipShift shift;
        if (l.active > 0 && h->multi && mip_shift_for(s.map, l.map, chunkGeom(h), srcOf(k), k, shift))
            h->run("mip", [&] {
                launch_mip_shifted(st, s.map, l.map, shift, l.active, l.list, s.col, s.nrm, l.col,
                                   l.qn, l.nrm);
            });
        else if (l.active > 0 && h->multi) h->run("mip", [&] {
            launch_chunk_mip(st, s.map, l.map, chunkGeom(h), srcOf(k), k, l.active, l.list, s.col,
                             s.nrm, l.col, l.qn, l.nrm);
        });
        else if (l.active > 0) h->run("mip", [&] {
            launch_mip(st, s.map, l.map, l.active, l.list, s.col, s.nrm, l.col, l.qn, l.nrm);
        });
    }
    // Measurement: the stored source voxels each filter read (timing runs,
    // one chunk).
    const bool countMip = (h->options & VOX_OPT_TIMING) && !h->replay && !h->multi;
    unsigned long long* mipCount = reinterpret_cast<unsigned long long*>(ctr + 256);
    if (countMip) {
        const FillSeg z[1] = {{ctr + 256, 2ull * VOX_MAX_LEVELS, 0u}};
        launch_fill_segs(st, z, 1);
        for (int k = 1; k < L; ++k) {
            if (h->lv[k].aliasOf >= 0) continue;
            LevelView s = view(h, srcOf(k));
            LevelView l = view(h, k);
            launch_mip_source_count(st, s.map, l.map, l.active, l.list, mipCount + k);
        }
    }
    // Node typing per depth (computeOctTreeNodes, Voxelizer.cu 931-1290).
    uint32_t border[3] = {h->extra[0], h->extra[1], h->extra[2]};
    std::vector<std::array<uint32_t, 3>> borders(L);
    // The samples of all depths in one launch, then each depth's comparison.
    static_assert(kMaxLevels >= VOX_MAX_LEVELS, "one sample entry per depth");
    SampleDepths S{};
    S.L = L;
    for (int a = 0; a < 3; ++a) S.B[a] = h->B[a];
    for (int d = L - 1; d >= 0; --d) {
        LevelView l = view(h, L - 1 - d);
        DepthBufs& t = h->tree[d];
        borders[d] = {border[0], border[1], border[2]};
        const uint32_t Ckyz[2] = {l.map.dims[1], l.map.dims[2]};
        const uint32_t byz[2] = {border[1], border[2]};
        const uint32_t fyz[2] = {t.Rd * h->B[1], t.Rd * h->B[2]};
        S.lm[d] = l.map;
        S.col[d] = l.col;
        S.cc[d] = t.cc.as<uint32_t>();
        S.flag[d] = h->nodeFlag.as<uint32_t>() + t.off;
        S.Rd[d] = t.Rd;
        for (int a = 0; a < 3; ++a) S.border[d][a] = border[a];
        S.xLimit[d] = type_x_limit(l.map.dims[0], border[0], t.Rd * h->B[0], Ckyz, byz, fyz);
        S.start[d] = t.off;
        for (int a = 0; a < 3; ++a) border[a] = std::max(border[a] >> 1, 1u);
    }
    S.start[L] = h->allNodes;
    const uint64_t bvox = (uint64_t)(h->B[0] + 2) * (h->B[1] + 2) * (h->B[2] + 2);
    uint32_t nb[VOX_MAX_LEVELS] = {0};
    if (h->multi) {
        // Several chunks: per-chunk samples and types OR-ed over chunks, and
        // bricks numbered in the order the brick writer writes them (the
        // chunk that completes a brick, inside one chunk before across
        // chunks, node order), by a key sort per depth.
        const ChunkGeom cg = chunkGeom(h);
        uint32_t* bcount = ctr + 64;
        h->run("fill", [&] {
            const FillSeg z[1] = {{bcount, (uint64_t)L, 0u}};
            launch_fill_segs(st, z, 1);
        });
        h->run("node_type", [&] {
            launch_chunk_node_type(st, S, cg);
            for (int d = L - 1; d >= 0; --d) {
                DepthBufs& t = h->tree[d];
                launch_chunk_brick_keys(st, S.flag[d], cg, L - 1 - d, t.Rd, h->B, S.border[d],
                                        h->keysA.as<unsigned long long>() + t.off,
                                        h->nodesA.as<uint32_t>() + t.off, bcount + d);
            }
        });
        occupancy();
        if (!h->check(hipMemcpyAsync(nb, bcount, 4 * L, hipMemcpyDeviceToHost, st), "brick counts") ||
            !h->check(hipStreamSynchronize(st), "node typing"))
            return 0;
        uint32_t rankBits = 0;
        while ((1ull << rankBits) < (uint64_t)h->nc[0] * h->nc[1] * h->nc[2]) ++rankBits;
        for (int d = 0; d < L; ++d) {
            DepthBufs& t = h->tree[d];
            t.bricks = nb[d];
            h->stats.bricks_total += nb[d];
            if (!t.brickNode.ensure((uint64_t)nb[d] * 4, h->err, "brick nodes") ||
                !t.bcol.ensure((uint64_t)nb[d] * bvox * 4, h->err, "brick colours") ||
                !t.bnrm.ensure((uint64_t)nb[d] * bvox * 4, h->err, "brick normals"))
                return 0;
            if (nb[d] > 1) {
                const int bits = (int)(3 * d + 1 + rankBits);
                size_t tb = 0;
                if (!h->check(chunk_sort_pairs(st, nullptr, &tb, nullptr, nullptr, nullptr, nullptr,
                                               nb[d], bits),
                              "sort size") ||
                    !h->sortTmp.ensure(tb, h->err, "sort scratch"))
                    return 0;
                tb = h->sortTmp.cap;
                if (!h->check(chunk_sort_pairs(st, h->sortTmp.p, &tb,
                                               h->keysA.as<unsigned long long>() + t.off,
                                               h->keysB.as<unsigned long long>(),
                                               h->nodesA.as<uint32_t>() + t.off,
                                               h->nodesB.as<uint32_t>(), nb[d], bits),
                              "brick order"))
                    return 0;
            }
            const uint32_t* sorted =
                nb[d] > 1 ? h->nodesB.as<uint32_t>() : h->nodesA.as<uint32_t>() + t.off;
            h->run("brick_assign", [&] {
                launch_chunk_brick_assign(st, t.nodes, S.flag[d], sorted, nb[d],
                                          t.types.as<uint8_t>(), t.bidx.as<uint32_t>(),
                                          t.brickNode.as<uint32_t>());
            });
        }
    } else {
        h->run("node_type", [&] {
            launch_node_sample_all(st, S);
            for (int d = L - 1; d >= 0; --d) {
                LevelView l = view(h, L - 1 - d);
                launch_node_compare(st, l.map, l.active, l.list, l.col, S.Rd[d], h->B, S.border[d],
                                    S.xLimit[d], S.cc[d], S.flag[d]);
            }
        });
        // One scan over the flags of every depth; a depth's brick addresses are
        // its scan values less the value at its first node.
        h->run("scan", [&] {
            scan(h, h->nodeFlag.as<uint32_t>(), h->nodeExcl.as<uint32_t>(), h->allNodes);
        });
        {
            uint64_t offs[VOX_MAX_LEVELS + 1];
            for (int d = 0; d < L; ++d) offs[d] = h->tree[d].off;
            offs[L] = h->allNodes;
            h->run("scan", [&] {
                launch_range_counts(st, h->nodeExcl.as<uint32_t>(), offs, L, ctr + 32);
            });
        }
        occupancy();
        if (h->replay) {
            std::memcpy(nb, h->planNb, sizeof(nb));
        } else {
            if (!h->check(hipMemcpyAsync(nb, ctr + 32, 4 * L, hipMemcpyDeviceToHost, st), "brick counts") ||
                !h->check(hipStreamSynchronize(st), "node typing"))
                return 0;
            std::memcpy(h->planNb, nb, sizeof(nb));
        }
        DepthOut dout{};
        dout.L = L;
        for (int d = 0; d < L; ++d) {
            DepthBufs& t = h->tree[d];
            t.bricks = nb[d];
            h->stats.bricks_total += nb[d];
            if (!t.brickNode.ensure((uint64_t)nb[d] * 4, h->err, "brick nodes") ||
                !t.bcol.ensure((uint64_t)nb[d] * bvox * 4, h->err, "brick colours") ||
                !t.bnrm.ensure((uint64_t)nb[d] * bvox * 4, h->err, "brick normals"))
                return 0;
            dout.bidx[d] = t.bidx.as<uint32_t>();
            dout.brickNode[d] = t.brickNode.as<uint32_t>();
            dout.types[d] = t.types.as<uint8_t>();
            dout.off[d] = t.off;
        }
        dout.off[L] = h->allNodes;
        h->run("brick_assign", [&] {
            launch_brick_assign_all(st, h->nodeFlag.as<uint32_t>(), h->nodeExcl.as<uint32_t>(), dout);
        });
    }
    // CollapseConstantBranches (Voxelizer.cu 701-774): the types are final.
    if (overlap) side.fork();
    for (int d = L - 2; d >= 0; --d) {
        h->run("collapse", [&] {
            launch_collapse(aux, h->tree[d].Rd, d + 1 == L - 1 ? 1 : 0,
                            h->tree[d].types.as<uint8_t>(), h->tree[d + 1].types.as<uint8_t>());
        });
    }
    // The bricks of all depths in one launch, the deepest (largest) first.
    GatherDepths GD{};
    for (int a = 0; a < 3; ++a) GD.B[a] = h->B[a];
    uint64_t nbAll = 0;
    for (int d = L - 1; d >= 0; --d) {
        if (nb[d] == 0) continue;
        LevelView l = view(h, L - 1 - d);
        DepthBufs& t = h->tree[d];
        const int e = GD.n++;
        GD.lm[e] = l.map;
        GD.col[e] = l.col;
        GD.qn[e] = l.qn;
        GD.brickNode[e] = t.brickNode.as<uint32_t>();
        GD.bcol[e] = t.bcol.as<uint32_t>();
        GD.bnrm[e] = t.bnrm.as<uint32_t>();
        GD.Rd[e] = t.Rd;
        for (int a = 0; a < 3; ++a) GD.border[e][a] = borders[d][a];
        GD.start[e] = nbAll;
        nbAll += nb[d];
    }
    GD.start[GD.n] = nbAll;
    if (nbAll >= (1ull << 31))
        return h->fail("more bricks than one launch's workgroups address"), 0;
    if (GD.n > 0) h->run("brick_gather", [&] { launch_brick_gather_all(st, GD); });
    // Compressed output: BC3 colours, BC1 normals (WriteCompressedBinaryBrick,
    // VoxelBrickWriter.cu 590-723, with the project's encoder).
    if (h->compressed) {
        const uint32_t e[3] = {h->B[0] + 2, h->B[1] + 2, h->B[2] + 2};
        const uint64_t blocks = bc_blocks_per_brick(e);
        for (int d = 0; d < L; ++d) {
            DepthBufs& t = h->tree[d];
            if (!t.ccol.ensure(t.bricks * blocks * 16, h->err, "compressed colours") ||
                !t.cnrm.ensure(t.bricks * blocks * 8, h->err, "compressed normals"))
                return 0;
            h->run("compress", [&] {
                launch_bc_bricks(st, t.bcol.as<uint32_t>(), t.bnrm.as<uint32_t>(), t.bricks, e,
                                 t.ccol.as<uint32_t>(), t.cnrm.as<uint32_t>());
            });
        }
    }
    side.join();
    h->workReleased = false;
    if (h->replay) {
        h->haveResult = true;
        h->prepared = false;
        return firstOverlap >= 0 ? 1 : -1;
    }
    (void)hipEventRecord(h->evStop, st);
    unsigned long long sat = 0;
    uint32_t scanErr = 0;
    if (!h->check(hipMemcpyAsync(&sat, satc, 8, hipMemcpyDeviceToHost, st), "counters") ||
        !h->check(hipMemcpyAsync(&scanErr, ctr + kScanErrWord, 4, hipMemcpyDeviceToHost, st),
                  "counters") ||
        !h->check(hipStreamSynchronize(st), "generate"))
        return 0;
    if (scanErr)
        return h->fail("a device-wide prefix sum gave up waiting on a predecessor; the result is "
                       "not valid (run vox_generate again)"), 0;
    h->planned = true;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, h->evStart, h->evStop);
    h->stats.ms_total = ms;
    h->stats.triangles = h->T;
    h->stats.pairs = h->pairs;
    h->stats.sat_tests = sat;
    h->stats.bins = h->numBins;
    h->stats.scan_values += h->numBins;   // the setup pass's bin scan
    if (countMip &&
        !h->check(hipMemcpy(h->stats.mip_source_voxels, mipCount, 8 * VOX_MAX_LEVELS,
                            hipMemcpyDeviceToHost),
                  "mip counts"))
        return 0;
    if (h->options & VOX_OPT_TIMING) collect_timings(h);
    h->haveResult = true;
    h->prepared = false;  // bins are consumed; a new generate re-runs setup
    if (out_dir) {
        if (!vox_write_files_impl(h, out_dir, h->p, h->dP, h->N, binary ? 1 : 0, h->extra, h->C,
                                  h->multi ? 1 : 0))
            return h->fail(std::string("failed to write GigaVoxels files to ") + out_dir), 0;
    }
    return firstOverlap >= 0 ? 1 : -1;
}


}  // namespace

int vox_generate(vox_ctx* h, const char* out_dir, int binary, int compressed) {
    if (!h || !h->valid || !idle(h)) return 0;
    h->compressed = compressed != 0;
    if (h->replay && out_dir) return h->fail("a replayed generate writes no files"), 0;
    if (h->replay && h->multi)
        return h->fail("a voxelizer with several chunks orders its bricks on the host's schedule "
                       "and is not captured in a graph"), 0;
    if (!h->prepared && !vox_compute_edges_face_normals_and_bounds(h)) return 0;
    if (!gen_front(h)) return 0;
    // Voxelize, colour and fill, group by group (Voxelizer.cu 2335-2553).
    for (int g = 0; g < (int)h->gargs.size(); ++g) {
        if (!h->gargs[g].overlaps) continue;
        gen_group_tiles(h, g);
        gen_voxelize(h, g);
        gen_group_tail(h, g);
    }
    return gen_back(h, out_dir, binary);
}

const char* vox_get_error(vox_ctx* h) {
    if (!h) return "null handle";
    if (!idle(h)) return "the handle is running vox_generate_async; call vox_wait first";
    h->errCache = h->err.str();
    return h->errCache.c_str();
}

int vox_reset(vox_ctx* h) {
    if (!h || !idle(h)) return 0;
    // resetCuda (Voxelizer.cu 324-344) frees, resets the device and
    // reallocates.  The device is not reset here: other users of the process
    // (torch, other handles) own memory on it.  Freeing everything and
    // recreating the stream and events leaves the handle in the same state.
    vox_deallocate_triangle_memory(h);
    vox_deallocate_voxel_memory(h);
    for (auto e : h->eventPool) (void)hipEventDestroy(e);
    h->eventPool.clear();
    h->eventNext = 0;
    if (h->evStart) (void)hipEventDestroy(h->evStart);
    if (h->evStop) (void)hipEventDestroy(h->evStop);
    h->evStart = h->evStop = nullptr;
    if (h->evFork) (void)hipEventDestroy(h->evFork);
    if (h->evJoin) (void)hipEventDestroy(h->evJoin);
    h->evFork = h->evJoin = nullptr;
    if (h->side) (void)hipStreamDestroy(h->side);
    h->side = nullptr;
    if (h->st) (void)hipStreamDestroy(h->st);
    h->st = nullptr;
    h->inited = false;
    if (!vox_init(h)) return 0;
    return vox_allocate_voxel_memory(h);
}

int vox_get_result(vox_ctx* h, vox_result* out) {
    if (!h || !h->valid || !out || !idle(h)) return 0;
    if (!h->haveResult) return h->fail("no result: call vox_generate first"), 0;
    std::memset(out, 0, sizeof(*out));
    out->levels = h->L;
    for (int a = 0; a < 3; ++a) {
        out->S[a] = h->S[a];
        out->brick_dim[a] = h->B[a];
    }
    const uint64_t nvox = (uint64_t)h->S[0] * h->S[1] * h->S[2];
    if (!(h->options & VOX_OPT_NO_OCCUPANCY)) {
        out->occupancy = h->bitmap.as<uint32_t>();
        out->occupancy_words = (nvox + 31) / 32;
    }
    for (int d = 0; d < h->L; ++d) {
        DepthBufs& t = h->tree[d];
        out->nodes[d] = t.nodes;
        out->types[d] = t.types.as<uint8_t>();
        out->const_colors[d] = t.cc.as<uint32_t>();
        out->brick_index[d] = t.bidx.as<uint32_t>();
        out->bricks[d] = t.bricks;
        out->brick_colors[d] = t.bcol.as<uint32_t>();
        out->brick_normals[d] = t.bnrm.as<uint32_t>();
        out->brick_node[d] = t.brickNode.as<uint32_t>();
    }
    return 1;
}

int vox_get_stats(vox_ctx* h, vox_stats* out) {
    if (!h || !out || !idle(h)) return 0;
    *out = h->stats;
    return 1;
}

int vox_copy_to_host(vox_ctx* h, void* dst, const void* src_device, uint64_t nbytes) {
    if (!h || !h->st || !idle(h)) return 0;
    if (nbytes == 0) return 1;
    if (!h->check(hipMemcpyAsync(dst, src_device, nbytes, hipMemcpyDeviceToHost, h->st), "copy") ||
        !h->check(hipStreamSynchronize(h->st), "copy"))
        return 0;
    return 1;
}

int vox_copy_level(vox_ctx* h, int32_t level, uint32_t* colors, float* normals, uint32_t dims[3]) {
    if (!h || !h->valid || !h->haveResult || !idle(h)) return 0;
    if (h->workReleased) return h->fail("vox_copy_level: the level tiles were released"), 0;
    if (level < 0 || level >= h->L) return h->fail("level out of range"), 0;
    LevelView l = view(h, level);
    // with several chunks level 0 is returned in the chunk-by-chunk extent
    // of the other levels (the chunks' voxels, zero past the voxel space)
    const bool pad = h->multi && level == 0;
    for (int a = 0; a < 3; ++a) dims[a] = pad ? h->nc[a] * h->C[a] : l.map.dims[a];
    if (!colors && !normals) return 1;
    if (!colors || !normals) return h->fail("vox_copy_level: colors and normals are both needed"), 0;
    const uint64_t n = (uint64_t)l.map.dims[0] * l.map.dims[1] * l.map.dims[2];
    DevBuf c, nr;
    if (!c.ensure(n * 4, h->err, "dense colours") || !nr.ensure(n * 12, h->err, "dense normals"))
        return 0;
    launch_dense_level(h->st, l.map, l.col, l.nrm, c.as<uint32_t>(), nr.as<float>());
    std::vector<uint32_t> hc;
    std::vector<float> hn;
    if (pad) {
        hc.resize(n);
        hn.resize(3 * n);
    }
    bool ok = h->check(hipMemcpyAsync(pad ? hc.data() : colors, c.p, n * 4, hipMemcpyDeviceToHost,
                                      h->st), "copy") &&
              h->check(hipMemcpyAsync(pad ? hn.data() : normals, nr.p, n * 12,
                                      hipMemcpyDeviceToHost, h->st), "copy") &&
              h->check(hipStreamSynchronize(h->st), "copy level");
    c.release();
    nr.release();
    if (ok && pad) {
        const uint32_t* sd = l.map.dims;
        const uint64_t out = (uint64_t)dims[0] * dims[1] * dims[2];
        std::fill(colors, colors + out, 0u);
        std::fill(normals, normals + 3 * out, 0.0f);
        const uint32_t ex = std::min(sd[0], dims[0]);
        for (uint32_t z = 0; z < std::min(sd[2], dims[2]); ++z)
            for (uint32_t y = 0; y < std::min(sd[1], dims[1]); ++y) {
                const uint64_t si = ((uint64_t)z * sd[1] + y) * sd[0];
                const uint64_t di = ((uint64_t)z * dims[1] + y) * dims[0];
                std::copy(hc.begin() + si, hc.begin() + si + ex, colors + di);
                std::copy(hn.begin() + 3 * si, hn.begin() + 3 * (si + ex), normals + 3 * di);
            }
    }
    return ok ? 1 : 0;
}

int vox_write_files_impl(vox_ctx* h, const char* out_dir, const float p[3], const float delta[3],
                         const uint32_t N[3], int binary, const uint32_t extra[3],
                         const uint32_t chunk[3], int multi);

int vox_get_compressed(vox_ctx* h, vox_compressed* out) {
    if (!h || !h->valid || !out || !idle(h)) return 0;
    if (!h->haveResult || !h->compressed) return 0;
    std::memset(out, 0, sizeof(*out));
    const uint32_t e[3] = {h->B[0] + 2, h->B[1] + 2, h->B[2] + 2};
    out->levels = h->L;
    out->blocks_per_brick = bc_blocks_per_brick(e);
    for (int d = 0; d < h->L; ++d) {
        out->bricks[d] = h->tree[d].bricks;
        out->colors[d] = h->tree[d].ccol.as<uint8_t>();
        out->normals[d] = h->tree[d].cnrm.as<uint8_t>();
    }
    return 1;
}

int vox_write_files(vox_ctx* h, const char* out_dir) {
    if (!h || !h->valid || !out_dir || !idle(h)) return 0;
    if (!vox_write_files_impl(h, out_dir, h->p, h->dP, h->N, 1, h->extra, h->C, h->multi ? 1 : 0))
        return h->fail(std::string("failed to write GigaVoxels files to ") + out_dir), 0;
    return 1;
}


// ---------------------------------------------------------------------------
// batch: many handles captured into one hipGraph (include/voxmi_batch.h)

namespace {

void buffer_list(vox_ctx* h, std::vector<void*>& out) {
    out.clear();
    DevBuf* top[] = {&h->verts, &h->fnorm, &h->bbox, &h->vn, &h->uv, &h->img, &h->grpFill,
                     &h->binCount, &h->binStart, &h->binTri, &h->scanTmp, &h->occ0, &h->hit0,
                     &h->sat, &h->work, &h->workCount, &h->rec, &h->recCount, &h->bitmap,
                     &h->counters, &h->nodeFlag, &h->nodeExcl, &h->fill0};
    for (DevBuf* b : top) out.push_back(b->p);
    for (auto& l : h->lv) {
        DevBuf* lb[] = {&l.mark, &l.excl, &l.tileIdx, &l.list, &l.col, &l.qn, &l.nrm};
        for (DevBuf* b : lb) out.push_back(b->p);
    }
    for (auto& t : h->tree) {
        DevBuf* tb[] = {&t.types, &t.cc, &t.bidx, &t.brickNode, &t.bcol, &t.bnrm, &t.ccol,
                        &t.cnrm};
        for (DevBuf* b : tb) out.push_back(b->p);
    }
}

// Host-side part of a job's signature: everything that sizes the pipeline
// besides the triangle data.
uint64_t host_signature(vox_ctx* h) {
    uint64_t z = 1469598103934665603ull;
    auto mixin = [&](const void* p, size_t n) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) z = (z ^ c[i]) * 1099511628211ull;
    };
    mixin(h->N, sizeof(h->N));
    mixin(h->B, sizeof(h->B));
    mixin(h->p, sizeof(h->p));
    mixin(h->dP, sizeof(h->dP));
    mixin(&h->options, sizeof(h->options));
    mixin(&h->compressed, sizeof(h->compressed));
    mixin(&h->persistentBlocks, sizeof(h->persistentBlocks));
    for (auto& g : h->groups) {
        const vox_group& d = g.desc;
        mixin(&d.ntri, sizeof(d.ntri));
        mixin(d.bbox_min, sizeof(d.bbox_min));
        mixin(d.bbox_max, sizeof(d.bbox_max));
        mixin(&d.is_terrain, sizeof(d.is_terrain));
        const int32_t flags[6] = {d.vnormals != nullptr, d.uvs != nullptr, d.rgba8 != nullptr,
                                  d.width, d.height, d.addr0 * 4 + d.addr1};
        mixin(flags, sizeof(flags));
    }
    return z;
}

// The job's input buffers as hash segments (vertices, vertex normals, uvs,
// images), each block of the hash launch covering at least 2048 words.
void input_segments(vox_ctx* h, uint32_t job, std::vector<HashSeg>& segs, uint32_t& blocks) {
    uint64_t nvn = 0, nuv = 0, nimg = 0;
    for (auto& g : h->groups) {
        if (g.desc.vnormals) nvn += g.desc.ntri;
        if (g.desc.uvs && g.desc.rgba8) {
            nuv += g.desc.ntri;
            nimg += (uint64_t)g.desc.width * g.desc.height * 4;
        }
    }
    const struct {
        const uint32_t* p;
        uint64_t n;
    } bufs[4] = {{h->verts.as<uint32_t>(), h->T * 9},
                 {h->vn.as<uint32_t>(), nvn * 9},
                 {h->uv.as<uint32_t>(), nuv * 6},
                 {h->img.as<uint32_t>(), nimg / 4}};
    for (uint64_t k = 0; k < 4; ++k) {
        if (bufs[k].n == 0 || !bufs[k].p) continue;
        uint64_t nb = (bufs[k].n + 2047) / 2048;
        if (nb > 1024) nb = 1024;
        segs.push_back(HashSeg{bufs[k].p, bufs[k].n, k + 1, job, blocks});
        blocks += (uint32_t)nb;
    }
}

}  // namespace

struct vox_batch {
    std::vector<vox_ctx*> jobs;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool valid = false, captured = false;
    std::vector<std::vector<void*>> bufs;
    std::vector<uint64_t> hostSig, devSig;
    DevBuf sums, segs;
    uint32_t nseg = 0, hashBlocks = 0;
    // one captured branch per job: its stream and the event joining it back
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> joined;
    hipEvent_t fork = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    // One persistent voxelize launch per wave of groups (wave w: the w-th
    // overlapping group of every job that has one) instead of one launch per
    // job and group: the job table, the per-job unit prefixes and a queue
    // word per wave, and the events that order the branches around it.
    // which jobs: 0 none, < 0 all, > 0 those with at most that many level-0
    // tiles (the others keep launches of their own)
    int32_t persistent = -1;
    unsigned voxBlocks = 0;   // 0: 4 workgroups per CU
    DevBuf jobTable, unitStart, queue;
    std::vector<hipEvent_t> preVox, postVox;
    vox_batch_stats stats{};
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
    // The segment table of every job's inputs (their buffers are fixed
    // between a capture and its launches).
    bool build_segments() {
        std::vector<HashSeg> v;
        uint32_t blocks = 0;
        for (size_t j = 0; j < jobs.size(); ++j) input_segments(jobs[j], (uint32_t)j, v, blocks);
        nseg = (uint32_t)v.size();
        hashBlocks = blocks;
        if (nseg == 0) return true;
        return segs.ensure(sizeof(HashSeg) * v.size(), err, "hash segments") &&
               check(hipMemcpy(segs.p, v.data(), sizeof(HashSeg) * v.size(), hipMemcpyHostToDevice),
                     "hash segments");
    }
    bool checksums(std::vector<uint64_t>& out) {
        launch_fill_u32(st, sums.as<uint32_t>(), 2 * jobs.size(), 0u);
        launch_hash_segs(st, segs.as<HashSeg>(), nseg, hashBlocks, sums.as<unsigned long long>());
        out.assign(jobs.size(), 0);
        return check(hipMemcpyAsync(out.data(), sums.p, 8 * jobs.size(), hipMemcpyDeviceToHost, st),
                     "checksum") &&
               check(hipStreamSynchronize(st), "checksum");
    }
    bool ensure_events(std::vector<hipEvent_t>& v, size_t n) {
        while (v.size() < n) {
            hipEvent_t e = nullptr;
            if (!check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"))
                return false;
            v.push_back(e);
        }
        return true;
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        captured = false;
    }
};

extern "C" {

vox_batch* vox_batch_create(vox_ctx* const* jobs, int32_t njobs) {
    vox_batch* b = new (std::nothrow) vox_batch;
    if (!b) return nullptr;
    if (!jobs || njobs <= 0) {
        b->fail("batch: no jobs");
        return b;
    }
    for (int32_t j = 0; j < njobs; ++j) {
        vox_ctx* h = jobs[j];
        if (!h || !h->valid || !idle(h) || !h->planned || !h->triMem) {
            b->fail("batch: job " + std::to_string(j) +
                    " has no eager vox_generate with uploaded triangles to size it");
            return b;
        }
        for (int32_t i = 0; i < j; ++i)
            if (jobs[i] == h) {
                b->fail("batch: a handle appears twice");
                return b;
            }
        b->jobs.push_back(h);
    }
    if (!b->check(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking), "hipStreamCreate") ||
        !b->sums.ensure(8 * b->jobs.size(), b->err, "checksums"))
        return b;
    for (auto& e : b->ev)
        if (!b->check(hipEventCreate(&e), "hipEventCreate")) return b;
    if (!b->check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate"))
        return b;
    b->branch.assign(b->jobs.size(), nullptr);
    b->joined.assign(b->jobs.size(), nullptr);
    for (size_t j = 0; j < b->jobs.size(); ++j)
        if (!b->check(hipStreamCreateWithFlags(&b->branch[j], hipStreamNonBlocking),
                      "hipStreamCreate") ||
            !b->check(hipEventCreateWithFlags(&b->joined[j], hipEventDisableTiming),
                      "hipEventCreate"))
            return b;
    b->stats.jobs = njobs;
    b->valid = true;
    return b;
}

void vox_batch_destroy(vox_batch* b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    b->drop_graph();
    for (auto e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto e : b->joined)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b->preVox)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b->postVox)
        if (e) (void)hipEventDestroy(e);
    b->jobTable.release();
    b->unitStart.release();
    b->queue.release();
    for (auto q : b->branch)
        if (q) (void)hipStreamDestroy(q);
    b->sums.release();
    b->segs.release();
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

const char* vox_batch_get_error(vox_batch* b) {
    if (!b) return "null batch";
    b->errCache = b->err.str();
    return b->errCache.c_str();
}

int vox_batch_capture(vox_batch* b) {
    if (!b || !b->valid) return 0;
    b->drop_graph();
    const auto t0 = std::chrono::steady_clock::now();
    for (vox_ctx* h : b->jobs)
        if (!idle(h)) return b->fail("batch: a job is running vox_generate_async"), 0;
    for (vox_ctx* h : b->jobs)
        if (!h->check(hipStreamSynchronize(h->st), "synchronize")) return b->fail(h->err.str()), 0;
    b->bufs.assign(b->jobs.size(), {});
    for (size_t j = 0; j < b->jobs.size(); ++j) buffer_list(b->jobs[j], b->bufs[j]);
    // Fork-join capture: every job is a branch of its own, forked from and
    // joined back into the batch stream, so the graph holds independent
    // chains the device can run side by side.  With the persistent voxelize
    // (the default) the chains meet once per wave of groups: every job's work
    // list of its w-th overlapping group, then one launch that voxelizes all
    // of them from a device queue, then every job's colour pass.
    const size_t J = b->jobs.size();
    std::vector<std::vector<int>> groupsOf(J);
    std::vector<char> inPers(J, 0);
    size_t waves = 0, entries = 0;
    for (size_t j = 0; j < J; ++j) {
        inPers[j] = b->persistent < 0 ||
                    (b->persistent > 0 && b->jobs[j]->planAct[0] <= (uint32_t)b->persistent);
        if (!inPers[j]) continue;
        for (size_t g = 0; g < b->jobs[j]->gargs.size(); ++g)
            if (b->jobs[j]->gargs[g].overlaps) groupsOf[j].push_back((int)g);
        waves = std::max(waves, groupsOf[j].size());
        entries += groupsOf[j].size();
    }
    // heaviest jobs first in every wave's table: their long units start early
    std::vector<size_t> order(J);
    for (size_t j = 0; j < J; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) {
        return b->jobs[a]->planPairs > b->jobs[c]->planPairs;
    });
    unsigned blocks = b->voxBlocks;
    if (blocks == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
            cus <= 0)
            cus = 256;
        blocks = 4u * (unsigned)cus;
    }
    const bool pers = entries > 0;
    if (pers && (!b->jobTable.ensure(sizeof(VoxUnitJob) * entries, b->err, "batch job table") ||
                 !b->unitStart.ensure(4 * (entries + waves), b->err, "batch unit offsets") ||
                 !b->queue.ensure(4 * waves, b->err, "batch queues") ||
                 !b->ensure_events(b->preVox, J) || !b->ensure_events(b->postVox, waves)))
        return 0;
    std::vector<VoxUnitJob> table;
    table.reserve(entries);
    if (!b->check(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal), "begin capture"))
        return 0;
    bool ok = b->check(hipEventRecord(b->fork, b->st), "fork");
    std::vector<hipStream_t> own(J, nullptr);
    for (size_t j = 0; j < J; ++j) {
        own[j] = b->jobs[j]->st;
        b->jobs[j]->st = b->branch[j];
        b->jobs[j]->replay = true;
        b->jobs[j]->prepared = false;
    }
    auto job_fail = [&](size_t j) {
        b->fail("batch: capture of job " + std::to_string(j) + " failed: " + b->jobs[j]->err.str());
        return false;
    };
    for (size_t j = 0; j < J && ok; ++j) {
        vox_ctx* h = b->jobs[j];
        ok = b->check(hipStreamWaitEvent(b->branch[j], b->fork, 0), "fork");
        if (!ok) break;
        if (!pers || !inPers[j]) {
            ok = (vox_compute_edges_face_normals_and_bounds(h) == 1 &&
                  vox_generate(h, nullptr, 1, h->compressed ? 1 : 0) != 0) || job_fail(j);
            continue;
        }
        ok = (vox_compute_edges_face_normals_and_bounds(h) == 1 && gen_front(h) != 0) ||
             job_fail(j);
        if (ok) {
            std::vector<int> now;
            for (size_t g = 0; g < h->gargs.size(); ++g)
                if (h->gargs[g].overlaps) now.push_back((int)g);
            if (now != groupsOf[j]) ok = b->fail("batch: job " + std::to_string(j) +
                                                 " changed its groups during the capture");
        }
    }
    for (size_t w = 0; w < waves && ok && pers; ++w) {
        const size_t first = table.size();
        for (size_t k = 0; k < J && ok; ++k) {
            const size_t j = order[k];
            if (groupsOf[j].size() <= w) continue;
            vox_ctx* h = b->jobs[j];
            gen_group_tiles(h, groupsOf[j][w]);
            table.push_back(gen_voxelize_job(h, groupsOf[j][w]));
            ok = b->check(hipEventRecord(b->preVox[j], b->branch[j]), "join") &&
                 b->check(hipStreamWaitEvent(b->st, b->preVox[j], 0), "join");
        }
        if (!ok) break;
        const uint32_t n = (uint32_t)(table.size() - first);
        launch_voxelize_batch(b->st, b->jobTable.as<VoxUnitJob>() + first, n,
                              b->unitStart.as<uint32_t>() + first + w, b->queue.as<uint32_t>() + w,
                              blocks);
        ok = b->check(hipGetLastError(), "batched voxelize") &&
             b->check(hipEventRecord(b->postVox[w], b->st), "fork");
        for (size_t j = 0; j < J && ok; ++j) {
            if (groupsOf[j].size() <= w) continue;
            ok = b->check(hipStreamWaitEvent(b->branch[j], b->postVox[w], 0), "fork");
            if (ok) gen_group_tail(b->jobs[j], groupsOf[j][w]);
        }
    }
    for (size_t j = 0; j < J && ok; ++j) {
        if (pers && inPers[j]) ok = gen_back(b->jobs[j], nullptr, 1) != 0 || job_fail(j);
        ok = ok && b->check(hipEventRecord(b->joined[j], b->branch[j]), "join") &&
             b->check(hipStreamWaitEvent(b->st, b->joined[j], 0), "join");
    }
    for (size_t j = 0; j < J; ++j) {
        b->jobs[j]->replay = false;
        b->jobs[j]->st = own[j];
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->st, &g);
    if (!ok || !b->check(e, "end capture")) {
        if (g) (void)hipGraphDestroy(g);
        return 0;
    }
    b->graph = g;
    if (!table.empty() &&
        !b->check(hipMemcpy(b->jobTable.p, table.data(), sizeof(VoxUnitJob) * table.size(),
                            hipMemcpyHostToDevice),
                  "batch job table")) {
        b->drop_graph();
        return 0;
    }
    {
        uint64_t own = 0;   // launches of the jobs outside the persistent launch
        for (size_t j = 0; j < J; ++j)
            if (!inPers[j] || !pers)
                for (const GroupArgs& ga : b->jobs[j]->gargs) own += ga.overlaps ? 1 : 0;
        b->stats.voxelize_launches = (uint32_t)(own + (pers ? waves : 0));
    }
    b->stats.voxelize_blocks = pers ? blocks : 0u;
    // the capture must not have moved any buffer (sizes are the planned ones)
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        std::vector<void*> now;
        buffer_list(b->jobs[j], now);
        if (now != b->bufs[j]) {
            b->drop_graph();
            return b->fail("batch: a buffer was reallocated during capture"), 0;
        }
    }
    size_t nn = 0;
    (void)hipGraphGetNodes(b->graph, nullptr, &nn);
    b->stats.graph_nodes = nn;
    if (!b->check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate")) {
        b->drop_graph();
        return 0;
    }
    b->hostSig.resize(b->jobs.size());
    for (size_t j = 0; j < b->jobs.size(); ++j) b->hostSig[j] = host_signature(b->jobs[j]);
    if (!b->build_segments() || !b->checksums(b->devSig)) return 0;
    b->captured = true;
    b->stats.ms_capture =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 1;
}

int vox_batch_launch(vox_batch* b) {
    if (!b || !b->valid) return 0;
    if (!b->captured) return b->fail("batch: capture first"), 0;
    for (vox_ctx* h : b->jobs)
        if (!idle(h)) return b->fail("batch: a job is running vox_generate_async"), 0;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        vox_ctx* h = b->jobs[j];
        std::vector<void*> now;
        buffer_list(h, now);
        if (now != b->bufs[j] || host_signature(h) != b->hostSig[j])
            return b->fail("batch: job " + std::to_string(j) +
                           " changed since the capture (buffers or parameters); capture again"), 0;
        if (!h->check(hipStreamSynchronize(h->st), "synchronize")) return 0;
    }
    (void)hipEventRecord(b->ev[0], b->st);
    std::vector<uint64_t> sig;
    if (!b->checksums(sig)) return 0;
    for (size_t j = 0; j < b->jobs.size(); ++j)
        if (sig[j] != b->devSig[j])
            return b->fail("batch: the triangles of job " + std::to_string(j) +
                           " changed since the capture; capture again"), 0;
    (void)hipEventRecord(b->ev[1], b->st);
    if (!b->check(hipGraphLaunch(b->exec, b->st), "graph launch")) return 0;
    (void)hipEventRecord(b->ev[2], b->st);
    if (!b->check(hipStreamSynchronize(b->st), "graph")) return 0;
    float a = 0, c = 0;
    (void)hipEventElapsedTime(&a, b->ev[0], b->ev[1]);
    (void)hipEventElapsedTime(&c, b->ev[1], b->ev[2]);
    b->stats.ms_check_last = a;
    b->stats.ms_graph_last = c;
    ++b->stats.launches;
    for (vox_ctx* h : b->jobs) h->haveResult = true;
    return 1;
}

int vox_batch_set_voxelize(vox_batch* b, int32_t persistent, uint32_t blocks) {
    if (!b || !b->valid) return 0;
    b->persistent = persistent;
    b->voxBlocks = blocks;
    b->drop_graph();   // takes effect at the next capture
    return 1;
}

int vox_batch_get_stats(vox_batch* b, vox_batch_stats* out) {
    if (!b || !b->valid || !out) return 0;
    *out = b->stats;
    return 1;
}

int vox_set_launch(vox_ctx* h, uint32_t persistent_blocks) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (persistent_blocks == 0 || persistent_blocks > (1u << 20))
        return h->fail("persistent_blocks out of range"), 0;
    h->persistentBlocks = persistent_blocks;
    return 1;
}

int vox_generate_async(vox_ctx* h, int32_t compressed) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (!h->inited && !vox_init(h)) return 0;
    if (h->worker.joinable()) h->worker.join();  // finished, never waited for
    h->async.store(1, std::memory_order_release);
    try {
        h->worker = std::thread([h, compressed] {
            t_worker = h;
            (void)hipSetDevice(0);
            const int r = vox_generate(h, nullptr, 1, compressed);
            h->asyncRet = r;
            t_worker = nullptr;
            h->async.store(2, std::memory_order_release);
        });
    } catch (...) {
        h->async.store(0);
        return h->fail("vox_generate_async: no worker thread"), 0;
    }
    return 1;
}

int vox_wait(vox_ctx* h, int32_t block) {
    if (!h) return 0;
    const int s = h->async.load(std::memory_order_acquire);
    if (s == 0) return h->fail("vox_wait: no vox_generate_async was started"), 0;
    if (s == 1 && !block) return 2;
    h->worker.join();
    h->async.store(0);
    return h->asyncRet;
}

void* vox_device_alloc(uint64_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();   // not left behind as the runtime's last error
        return nullptr;
    }
    return p;
}

int vox_device_upload(void* dst, uint64_t offset, const void* src, uint64_t bytes) {
    if (!dst || (bytes && !src)) return 0;
    if (bytes == 0) return 1;
    return hipMemcpy(static_cast<uint8_t*>(dst) + offset, src, bytes, hipMemcpyHostToDevice) ==
                   hipSuccess
               ? 1
               : 0;
}

int vox_device_download(void* dst, const void* src, uint64_t offset, uint64_t bytes) {
    if (!src || (bytes && !dst)) return 0;
    if (bytes == 0) return 1;
    return hipMemcpy(dst, static_cast<const uint8_t*>(src) + offset, bytes, hipMemcpyDeviceToHost) ==
                   hipSuccess
               ? 1
               : 0;
}

int vox_device_copy(void* dst, const void* src, uint64_t bytes) {
    if (!dst || !src || (bytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u)) return 0;
    launch_copy_u128(nullptr, src, dst, bytes / 16);
    return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 1 : 0;
}

int vox_device_copy_scattered(void* dst, const void* src, uint64_t bytes) {
    if (!dst || !src || bytes < 256 || (bytes & (bytes - 1)) || ((uintptr_t)dst & 15u) ||
        ((uintptr_t)src & 15u))
        return 0;
    launch_copy_scattered(nullptr, src, dst, bytes / 256);
    return hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess ? 1 : 0;
}

int vox_device_free(void* p) {
    if (!p) return 1;
    if (hipDeviceSynchronize() != hipSuccess) return 0;
    return hipFree(p) == hipSuccess ? 1 : 0;
}

int vox_release_work_memory(vox_ctx* h) {
    if (!h || !h->valid || !idle(h)) return 0;
    if (h->st && !h->check(hipStreamSynchronize(h->st), "synchronize")) return 0;
    // only this handle's stream uses these buffers, and it is idle now: they
    // go to the block cache (no hipFree, which would wait for the device)
    for (auto& l : h->lv) {
        l.col.recycle(); l.qn.recycle(); l.nrm.recycle();
        l.active = 0;
    }
    h->occ0.recycle(); h->hit0.recycle(); h->fill0.recycle(); h->work.recycle();
    h->workCount.recycle();
    h->rec.recycle(); h->recCount.recycle(); h->binTri.recycle(); h->binCount.recycle();
    h->binStart.recycle(); h->scanTmp.recycle();
    h->verts.recycle(); h->fnorm.recycle(); h->bbox.recycle(); h->vn.recycle();
    h->gbounds.recycle(); h->uv.recycle(); h->img.recycle(); h->grpFill.recycle();
    vox_deallocate_triangle_memory(h);
    h->planned = false;
    h->workReleased = true;
    return 1;
}

}  // extern "C" (batch)

}  // extern "C"
