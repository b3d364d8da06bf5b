// vox_server.hip -- device-resident brick server: request admission into the
// brick pool with LRU eviction (C ABI in include/voxmi_serve.h).
//
// The reference runs this on host threads: node-usage lists are read back,
// deduplicated in a std::map, bricks the GPU lacks go on an upload list and
// BrickPool::update copies each into the front slot of an LRU list, moving
// used slots to its back (GigaVoxelsOctTree.cpp 430-652,
// GigaVoxelsBrickPool.cpp 406-427, 561-682).  Here the whole step stays on
// the device and is deterministic: touched slots carry the frame number, new
// pending nodes join a FIFO in ascending texel order, and victims are the
// slots with the oldest use (empty first, ties by slot index), found with a
// 257-bucket age histogram instead of a linked list.
#include <hip/hip_runtime.h>

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

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kResidentBit = 0x80000000u;
constexpr int kB = 256;
constexpr uint32_t kKeys = 257;  // 0 empty, 1..255 by age (older first), 256 used now

struct DepthTabs {
    const uint8_t* types[kMaxLevels];
    const uint32_t* bidx[kMaxLevels];
    const uint32_t* texel[kMaxLevels];
};

unsigned grid_for(uint64_t n) {
    const uint64_t g = (n + kB - 1) / kB;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

// texel -> (depth, node) of every node in the tree.
__global__ void k_srv_info(uint64_t n, uint32_t sh, uint32_t d, const uint32_t* __restrict__ texel,
                           uint32_t* __restrict__ info) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t = texel[i];
        if (t != kNone) info[t] = (d << 27) | (uint32_t)i;
    }
}

// Slots already holding a brick (the root's, uploaded when the octree was
// built): recovered from the resident bit and brick pointer of the texels.
__global__ void k_srv_adopt(uint64_t ntex, const uint32_t* __restrict__ tex, uint32_t ex,
                            uint32_t ey, uint32_t ez, uint32_t sx, uint32_t sy, uint32_t nslots,
                            uint32_t* __restrict__ nodeSlot, uint32_t* __restrict__ slotNode) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntex;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w0 = tex[2 * t], w1 = tex[2 * t + 1];
        if (!(w0 & kResidentBit)) continue;
        const uint32_t x = (((w1 >> 20) & 1023u) - 1u) / ex;
        const uint32_t y = (((w1 >> 10) & 1023u) - 1u) / ey;
        const uint32_t z = ((w1 & 1023u) - 1u) / ez;
        const uint32_t s = (z * sy + y) * sx + x;
        if (s < nslots) {
            nodeSlot[t] = s;
            slotNode[s] = (uint32_t)t;
        }
    }
}

__device__ __forceinline__ uint32_t parent_texel(uint32_t info, const DepthTabs& D) {
    const uint32_t d = info >> 27;
    if (d == 0) return kNone;
    const uint32_t n = info & 0x07ffffffu;
    const uint32_t sh = d;  // 2^d nodes per axis
    const uint32_t m = (1u << sh) - 1u;
    const uint32_t x = n & m, y = (n >> sh) & m, z = n >> (2 * sh);
    const uint32_t p = (((z >> 1) << (2 * (sh - 1))) | ((y >> 1) << (sh - 1)) | (x >> 1));
    return D.texel[d - 1][p];
}

__device__ __forceinline__ bool non_const(uint32_t info, const DepthTabs& D) {
    return D.types[info >> 27][info & 0x07ffffffu] == 1;
}

__device__ __forceinline__ void mark_pending(uint32_t t, unsigned long long emit,
                                             uint32_t* __restrict__ pend,
                                             unsigned long long* __restrict__ pendEmit) {
    if (atomicCAS(&pend[t], 0u, 2u) == 0u) pendEmit[t] = emit;
}

// 1. touch used slots, mark new pending nodes (and their parents).
__global__ void k_srv_touch(const uint32_t* __restrict__ ids, const uint64_t* __restrict__ times,
                            uint32_t nids, uint32_t ndx, uint32_t ndy, uint32_t ndz, uint32_t frame,
                            DepthTabs D, const uint32_t* __restrict__ info,
                            const uint32_t* __restrict__ nodeSlot, uint32_t* __restrict__ slotLast,
                            uint32_t* __restrict__ pend,
                            unsigned long long* __restrict__ pendEmit) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nids; k += gridDim.x * blockDim.x) {
        const uint32_t id = ids[k];
        const uint32_t x = ((id >> 20) & 1023u) - 1u, y = ((id >> 10) & 1023u) - 1u,
                       z = (id & 1023u) - 1u;
        if (x >= ndx || y >= ndy || z >= ndz) continue;
        const uint32_t t = (z * ndy + y) * ndx + x;
        const uint32_t inf = info[t];
        if (inf == kNone) continue;  // stale id outside the tree
        const uint32_t s = nodeSlot[t];
        if (s != kNone) {
            slotLast[s] = frame;
            continue;
        }
        if (!non_const(inf, D)) continue;
        const unsigned long long emit = (unsigned long long)times[k];
        mark_pending(t, emit, pend, pendEmit);
        const uint32_t p = parent_texel(inf, D);
        if (p != kNone && nodeSlot[p] == kNone && non_const(info[p], D))
            mark_pending(p, emit, pend, pendEmit);
    }
}

__global__ void k_srv_newflags(uint64_t ntex, const uint32_t* __restrict__ pend,
                               uint32_t* __restrict__ flag) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntex;
         t += (uint64_t)gridDim.x * blockDim.x)
        flag[t] = pend[t] == 2u ? 1u : 0u;
}

// 2. append the new pending texels to the FIFO in ascending order.
// ctr: [0] head, [1] tail (monotonic), [2] admitted this frame, [3] evictions
__global__ void k_srv_enqueue(uint64_t ntex, const uint32_t* __restrict__ flag,
                              const uint32_t* __restrict__ excl, uint32_t* __restrict__ pend,
                              const unsigned long long* __restrict__ pendEmit,
                              uint32_t* __restrict__ ring, unsigned long long* __restrict__ ringEmit,
                              uint32_t cap, const uint32_t* __restrict__ ctr) {
    const uint32_t tail = ctr[1];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntex;
         t += (uint64_t)gridDim.x * blockDim.x) {
        if (!flag[t]) continue;
        const uint32_t pos = (tail + excl[t]) % cap;
        ring[pos] = (uint32_t)t;
        ringEmit[pos] = pendEmit[t];
        pend[t] = 1u;
    }
}

__global__ void k_srv_tail(const uint32_t* __restrict__ excl, uint64_t ntex,
                           uint32_t* __restrict__ ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctr[1] += excl[ntex];
}

// 3a. slot keys and their histogram.
__global__ void k_srv_keys(uint32_t nslots, uint32_t frame, const uint32_t* __restrict__ slotNode,
                           const uint32_t* __restrict__ slotLast, uint32_t* __restrict__ key,
                           uint32_t* __restrict__ hist) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x) {
        uint32_t k;
        if (slotNode[s] == kNone) {
            k = 0;
        } else {
            const uint32_t age = frame - slotLast[s];
            k = age == 0 ? 256u : 256u - (age < 255u ? age : 255u);
        }
        key[s] = k;
        atomicAdd(&hist[k], 1u);
    }
}

// 3b. how many to admit and the key threshold.  sel: [0] n, [1] key T,
// [2] slots to take with key T.
__global__ void k_srv_threshold(const uint32_t* __restrict__ hist, const uint32_t* __restrict__ ctr,
                                uint32_t maxUp, uint32_t* __restrict__ sel) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t pending = ctr[1] - ctr[0];
    uint32_t avail = 0;
    for (uint32_t k = 0; k < 256; ++k) avail += hist[k];
    uint32_t n = pending < maxUp ? pending : maxUp;
    n = n < avail ? n : avail;
    uint32_t cum = 0, T = 0, takeEq = 0;
    if (n > 0) {
        for (T = 0; T < 256; ++T) {
            if (cum + hist[T] >= n) break;
            cum += hist[T];
        }
        takeEq = n - cum;
    }
    sel[0] = n;
    sel[1] = T;
    sel[2] = takeEq;
}

__global__ void k_srv_eqflags(uint32_t nslots, const uint32_t* __restrict__ key,
                              const uint32_t* __restrict__ sel, uint32_t* __restrict__ flag) {
    const uint32_t T = sel[1];
    const bool any = sel[0] > 0;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x)
        flag[s] = (any && key[s] == T) ? 1u : 0u;
}

__global__ void k_srv_selflags(uint32_t nslots, const uint32_t* __restrict__ key,
                               const uint32_t* __restrict__ sel, const uint32_t* __restrict__ eqExcl,
                               uint32_t* __restrict__ flag) {
    const uint32_t n = sel[0], T = sel[1], takeEq = sel[2];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x) {
        const uint32_t k = key[s];
        flag[s] = (n > 0 && (k < T || (k == T && eqExcl[s] < takeEq))) ? 1u : 0u;
    }
}

// 4. admission: FIFO entry i -> i-th selected slot (ascending slot index).
__global__ void k_srv_admit(uint32_t nslots, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ excl, const uint32_t* __restrict__ sel,
                            uint32_t frame, DepthTabs D, const uint32_t* __restrict__ info,
                            uint32_t* __restrict__ ring, const unsigned long long* __restrict__ ringEmit,
                            uint32_t cap, uint32_t* __restrict__ ctr, uint32_t* __restrict__ nodeSlot,
                            uint32_t* __restrict__ slotNode, uint32_t* __restrict__ slotLast,
                            uint32_t* __restrict__ pend, uint32_t* __restrict__ tex,
                            uint4* __restrict__ req, unsigned long long* __restrict__ reqEmit,
                            uint32_t* __restrict__ nreq) {
    const uint32_t n = sel[0];
    const uint32_t head = ctr[0];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x) {
        if (!flag[s]) continue;
        const uint32_t i = excl[s];
        if (i >= n) continue;
        const uint32_t pos = (head + i) % cap;
        const uint32_t t = ring[pos];
        const uint32_t old = slotNode[s];
        if (old != kNone) {
            nodeSlot[old] = kNone;
            tex[2 * (uint64_t)old] &= ~kResidentBit;
            atomicAdd(&ctr[3], 1u);
        }
        slotNode[s] = t;
        slotLast[s] = frame;
        nodeSlot[t] = s;
        pend[t] = 0u;
        const uint32_t inf = info[t];
        const uint32_t d = inf >> 27, node = inf & 0x07ffffffu;
        req[i] = make_uint4(d, D.bidx[d][node], s, t);
        reqEmit[i] = ringEmit[pos];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *nreq = n;
}

__global__ void k_srv_advance(const uint32_t* __restrict__ sel, uint32_t* __restrict__ ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr[0] += sel[0];
        ctr[2] = sel[0];
    }
}

// Latency of each admitted request: copy done - emitted.
__global__ void k_srv_latency(const uint32_t* __restrict__ nreq,
                              const unsigned long long* __restrict__ reqEmit,
                              const unsigned long long* __restrict__ done,
                              unsigned long long* __restrict__ log, uint64_t logCap,
                              unsigned long long* __restrict__ logCount) {
    const uint32_t n = *nreq;
    const unsigned long long base = *logCount;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long k = base + i;
        if (k < logCap) log[k] = done[i] - reqEmit[i];
    }
}

__global__ void k_srv_logcount(const uint32_t* __restrict__ nreq,
                               unsigned long long* __restrict__ logCount) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *logCount += *nreq;
}

}  // namespace

struct vox_server {
    vox_octree* tree = nullptr;
    bool valid = false;
    uint32_t maxUp = 0, nslots = 0, frame = 0;
    uint64_t ntex = 0;
    uint32_t ndim[3] = {0, 0, 0};
    hipStream_t st = nullptr;
    DepthTabs D{};
    DevBuf info, nodeSlot, slotNode, slotLast, pend, pendEmit, ring, ringEmit, ctr;
    DevBuf flagT, exclT, scanTmpT, key, hist, sel, flagS, eqExclS, exclS, scanTmpS;
    DevBuf req, reqEmit, nreq, done, log, logCount;
    uint64_t logCap = 1u << 22;
    vox_server_stats stats{};
    hipEvent_t ev[2] = {nullptr, nullptr};
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
void launch_brick_upload_timed(hipStream_t st, const BrickSrc& src, const uint4* req,
                               const uint32_t* nreq, uint32_t maxReq, const uint32_t e[3],
                               const uint32_t slots[3], const uint32_t dims[3], uint32_t* col,
                               uint32_t* nrm, uint32_t* tex, unsigned long long* done);
}

extern "C" {

vox_server* vox_server_create(vox_octree* tree, uint32_t max_uploads_per_frame) {
    vox_server* s = new (std::nothrow) vox_server;
    if (!s) return nullptr;
    s->tree = tree;
    OctreeRefs o;
    if (!tree || !octree_refs(tree, o)) {
        s->fail("server: invalid octree");
        return s;
    }
    if (o.levels > 9) {
        s->fail("server: more than 9 octree depths do not fit the texel table");
        return s;
    }
    s->st = o.st;
    for (int a = 0; a < 3; ++a) s->ndim[a] = o.ndim[a];
    s->ntex = (uint64_t)o.ndim[0] * o.ndim[1] * o.ndim[2];
    s->nslots = o.poolSlots[0] * o.poolSlots[1] * o.poolSlots[2];
    s->maxUp = max_uploads_per_frame ? max_uploads_per_frame : s->nslots;
    if (s->maxUp > s->nslots) s->maxUp = s->nslots;
    for (int d = 0; d < kMaxLevels; ++d) {
        s->D.types[d] = o.types[d];
        s->D.bidx[d] = o.bidx[d];
        s->D.texel[d] = o.texel[d];
    }
    const uint64_t nt = s->ntex, ns = s->nslots;
    if (!s->info.ensure(4 * nt, s->err, "info") || !s->nodeSlot.ensure(4 * nt, s->err, "nodeSlot") ||
        !s->slotNode.ensure(4 * ns, s->err, "slotNode") ||
        !s->slotLast.ensure(4 * ns, s->err, "slotLast") || !s->pend.ensure(4 * nt, s->err, "pend") ||
        !s->pendEmit.ensure(8 * nt, s->err, "pendEmit") || !s->ring.ensure(4 * nt, s->err, "ring") ||
        !s->ringEmit.ensure(8 * nt, s->err, "ringEmit") || !s->ctr.ensure(16, s->err, "ctr") ||
        !s->flagT.ensure(4 * nt, s->err, "flagT") || !s->exclT.ensure(4 * (nt + 1), s->err, "exclT") ||
        !s->scanTmpT.ensure(4 * scan_tmp_words(nt), s->err, "scanTmpT") ||
        !s->key.ensure(4 * ns, s->err, "key") || !s->hist.ensure(4 * kKeys, s->err, "hist") ||
        !s->sel.ensure(16, s->err, "sel") || !s->flagS.ensure(4 * ns, s->err, "flagS") ||
        !s->eqExclS.ensure(4 * (ns + 1), s->err, "eqExclS") ||
        !s->exclS.ensure(4 * (ns + 1), s->err, "exclS") ||
        !s->scanTmpS.ensure(4 * scan_tmp_words(ns), s->err, "scanTmpS") ||
        !s->req.ensure(16ull * s->maxUp, s->err, "req") ||
        !s->reqEmit.ensure(8ull * s->maxUp, s->err, "reqEmit") || !s->nreq.ensure(4, s->err, "nreq") ||
        !s->done.ensure(8ull * s->maxUp, s->err, "done") ||
        !s->log.ensure(8 * s->logCap, s->err, "latency log") ||
        !s->logCount.ensure(8, s->err, "latency count"))
        return s;
    hipStream_t st = s->st;
    launch_fill_u32(st, s->info.as<uint32_t>(), nt, kNone);
    for (int d = 0; d < o.levels; ++d) {
        const uint64_t n = 1ull << (3 * d);
        k_srv_info<<<grid_for(n), kB, 0, st>>>(n, (uint32_t)d, (uint32_t)d, o.texel[d],
                                               s->info.as<uint32_t>());
    }
    launch_fill_u32(st, s->nodeSlot.as<uint32_t>(), nt, kNone);
    launch_fill_u32(st, s->slotNode.as<uint32_t>(), ns, kNone);
    launch_fill_u32(st, s->slotLast.as<uint32_t>(), ns, 0u);
    launch_fill_u32(st, s->pend.as<uint32_t>(), nt, 0u);
    launch_fill_u32(st, s->ctr.as<uint32_t>(), 4, 0u);
    launch_fill_u32(st, s->logCount.as<uint32_t>(), 2, 0u);
    k_srv_adopt<<<grid_for(nt), kB, 0, st>>>(nt, o.tex, o.B[0] + 2, o.B[1] + 2, o.B[2] + 2,
                                             o.poolSlots[0], o.poolSlots[1], s->nslots,
                                             s->nodeSlot.as<uint32_t>(), s->slotNode.as<uint32_t>());
    for (auto& e : s->ev)
        if (!s->check(hipEventCreate(&e), "hipEventCreate")) return s;
    int khz = 0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (!s->check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev),
                  "wall clock rate"))
        return s;
    s->stats.clock_khz = khz;
    if (!s->check(hipGetLastError(), "server init") ||
        !s->check(hipStreamSynchronize(st), "synchronize"))
        return s;
    s->valid = true;
    return s;
}

void vox_server_destroy(vox_server* s) {
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    for (auto e : s->ev)
        if (e) (void)hipEventDestroy(e);
    DevBuf* bufs[] = {&s->info, &s->nodeSlot, &s->slotNode, &s->slotLast, &s->pend, &s->pendEmit,
                      &s->ring, &s->ringEmit, &s->ctr, &s->flagT, &s->exclT, &s->scanTmpT,
                      &s->key, &s->hist, &s->sel, &s->flagS, &s->eqExclS, &s->exclS,
                      &s->scanTmpS, &s->req, &s->reqEmit, &s->nreq, &s->done, &s->log,
                      &s->logCount};
    for (DevBuf* b : bufs) b->release();
    delete s;
}

const char* vox_server_get_error(vox_server* s) {
    if (!s) return "null server";
    s->errCache = s->err.str();
    return s->errCache.c_str();
}

int vox_server_frame(vox_server* s, vox_raycaster* rc) {
    if (!s || !s->valid) return 0;
    const uint32_t* ids = nullptr;
    const uint64_t* times = nullptr;
    uint32_t nids = 0;
    if (!raycaster_requests(rc, &ids, &times, &nids))
        return s->fail("server: invalid ray caster"), 0;
    OctreeRefs o;
    if (!octree_refs(s->tree, o)) return s->fail("server: octree unavailable"), 0;
    hipStream_t st = s->st;
    const uint32_t frame = ++s->frame;
    const uint64_t nt = s->ntex, ns = s->nslots;
    (void)hipEventRecord(s->ev[0], st);
    if (nids)
        k_srv_touch<<<grid_for(nids), kB, 0, st>>>(
            ids, times, nids, s->ndim[0], s->ndim[1], s->ndim[2], frame, s->D,
            s->info.as<uint32_t>(), s->nodeSlot.as<uint32_t>(), s->slotLast.as<uint32_t>(),
            s->pend.as<uint32_t>(), s->pendEmit.as<unsigned long long>());
    k_srv_newflags<<<grid_for(nt), kB, 0, st>>>(nt, s->pend.as<uint32_t>(), s->flagT.as<uint32_t>());
    launch_scan(st, s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), nt, s->scanTmpT.as<uint32_t>());
    k_srv_enqueue<<<grid_for(nt), kB, 0, st>>>(
        nt, s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), s->pend.as<uint32_t>(),
        s->pendEmit.as<unsigned long long>(), s->ring.as<uint32_t>(),
        s->ringEmit.as<unsigned long long>(), (uint32_t)nt, s->ctr.as<uint32_t>());
    k_srv_tail<<<1, 64, 0, st>>>(s->exclT.as<uint32_t>(), nt, s->ctr.as<uint32_t>());
    launch_fill_u32(st, s->hist.as<uint32_t>(), kKeys, 0u);
    k_srv_keys<<<grid_for(ns), kB, 0, st>>>((uint32_t)ns, frame, s->slotNode.as<uint32_t>(),
                                             s->slotLast.as<uint32_t>(), s->key.as<uint32_t>(),
                                             s->hist.as<uint32_t>());
    k_srv_threshold<<<1, 64, 0, st>>>(s->hist.as<uint32_t>(), s->ctr.as<uint32_t>(), s->maxUp,
                                      s->sel.as<uint32_t>());
    k_srv_eqflags<<<grid_for(ns), kB, 0, st>>>((uint32_t)ns, s->key.as<uint32_t>(),
                                                s->sel.as<uint32_t>(), s->flagS.as<uint32_t>());
    launch_scan(st, s->flagS.as<uint32_t>(), s->eqExclS.as<uint32_t>(), ns,
                s->scanTmpS.as<uint32_t>());
    k_srv_selflags<<<grid_for(ns), kB, 0, st>>>((uint32_t)ns, s->key.as<uint32_t>(),
                                                 s->sel.as<uint32_t>(), s->eqExclS.as<uint32_t>(),
                                                 s->flagS.as<uint32_t>());
    launch_scan(st, s->flagS.as<uint32_t>(), s->exclS.as<uint32_t>(), ns,
                s->scanTmpS.as<uint32_t>());
    k_srv_admit<<<grid_for(ns), kB, 0, st>>>(
        (uint32_t)ns, s->flagS.as<uint32_t>(), s->exclS.as<uint32_t>(), s->sel.as<uint32_t>(),
        frame, s->D, s->info.as<uint32_t>(), s->ring.as<uint32_t>(),
        s->ringEmit.as<unsigned long long>(), (uint32_t)nt, s->ctr.as<uint32_t>(),
        s->nodeSlot.as<uint32_t>(), s->slotNode.as<uint32_t>(), s->slotLast.as<uint32_t>(),
        s->pend.as<uint32_t>(), o.tex, s->req.as<uint4>(), s->reqEmit.as<unsigned long long>(),
        s->nreq.as<uint32_t>());
    k_srv_advance<<<1, 64, 0, st>>>(s->sel.as<uint32_t>(), s->ctr.as<uint32_t>());
    const uint32_t e[3] = {o.B[0] + 2, o.B[1] + 2, o.B[2] + 2};
    launch_brick_upload_timed(st, o.src, s->req.as<uint4>(), s->nreq.as<uint32_t>(), s->maxUp, e,
                              o.poolSlots, o.poolDims, o.poolCol, o.poolNrm, o.tex,
                              s->done.as<unsigned long long>());
    k_srv_latency<<<grid_for(s->maxUp), kB, 0, st>>>(
        s->nreq.as<uint32_t>(), s->reqEmit.as<unsigned long long>(),
        s->done.as<unsigned long long>(), s->log.as<unsigned long long>(), s->logCap,
        s->logCount.as<unsigned long long>());
    k_srv_logcount<<<1, 64, 0, st>>>(s->nreq.as<uint32_t>(), s->logCount.as<unsigned long long>());
    (void)hipEventRecord(s->ev[1], st);
    uint32_t ctr[4] = {0, 0, 0, 0};
    if (!s->check(hipGetLastError(), "server kernels") ||
        !s->check(hipMemcpyAsync(ctr, s->ctr.p, 16, hipMemcpyDeviceToHost, st), "counters") ||
        !s->check(hipStreamSynchronize(st), "synchronize"))
        return 0;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, s->ev[0], s->ev[1]);
    s->stats.frames = frame;
    s->stats.requests += nids;
    s->stats.uploads += ctr[2];
    s->stats.evictions = ctr[3];
    s->stats.pending = ctr[1] - ctr[0];
    s->stats.last_uploads = ctr[2];
    s->stats.ms_update_last = ms;
    s->stats.ms_update_total += ms;
    return 1;
}

int vox_server_get_stats(vox_server* s, vox_server_stats* out) {
    if (!s || !s->valid || !out) return 0;
    *out = s->stats;
    // resident slots
    std::vector<uint32_t> sn(s->nslots);
    if (hipMemcpy(sn.data(), s->slotNode.p, 4ull * s->nslots, hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    uint32_t r = 0;
    for (uint32_t v : sn) r += v != kNone;
    out->resident = r;
    return 1;
}

int vox_server_latencies(vox_server* s, uint64_t* out, uint64_t cap, uint64_t* n) {
    if (!s || !s->valid || !n) return 0;
    unsigned long long cnt = 0;
    if (hipMemcpy(&cnt, s->logCount.p, 8, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    *n = cnt;
    uint64_t m = cnt < s->logCap ? cnt : s->logCap;
    if (m > cap) m = cap;
    if (m && out &&
        hipMemcpy(out, s->log.p, 8 * m, hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    return 1;
}

int vox_server_state(vox_server* s, const uint32_t** node_slot, const uint32_t** slot_node,
                     uint64_t* ntexels, uint32_t* nslots) {
    if (!s || !s->valid) return 0;
    if (node_slot) *node_slot = s->nodeSlot.as<uint32_t>();
    if (slot_node) *slot_node = s->slotNode.as<uint32_t>();
    if (ntexels) *ntexels = s->ntex;
    if (nslots) *nslots = s->nslots;
    return 1;
}

}  // extern "C"
