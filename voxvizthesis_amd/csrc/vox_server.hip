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
//
// One server can serve several octrees that share its brick pool, as the
// reference's BrickPool singleton serves every loaded tree of the pager
// (GigaVoxelsRenderer.cpp 1520-1582): a node is then named by its global
// texel, tree slot * texels per tree + texel, and the device table of
// attached trees gives each tree's node arrays, node texture and bricks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
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

// One attached tree as the kernels see it.
struct TreeTab {
    const uint8_t* types[kMaxLevels];
    const uint32_t* bidx[kMaxLevels];
    const uint32_t* texel[kMaxLevels];
    const uint32_t* col[kMaxLevels];   // the voxelizer's bricks
    const uint32_t* nrm[kMaxLevels];
    uint32_t* tex;                     // node texture, 2 words per texel
    uint32_t ndim[3];
    uint32_t live;
};

unsigned grid_for(uint64_t n) {
    const uint64_t g = (n + kB - 1) / kB;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

__device__ __forceinline__ uint32_t pack10(uint32_t x, uint32_t y, uint32_t z) {
    return (x << 20) | (y << 10) | z;
}

__global__ void k_srv_set_tab(TreeTab* __restrict__ tabs, uint32_t i, TreeTab v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) tabs[i] = v;
}

// texel -> (depth, node) of every node in the tree.
__global__ void k_srv_info(uint64_t n, uint32_t d, const uint32_t* __restrict__ texel,
                           uint32_t* __restrict__ info) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t = texel[i];
        if (t != kNone) info[t] = (d << 27) | (uint32_t)i;
    }
}

// Slots already holding a brick (a root uploaded when the octree was built):
// recovered from the resident bit and brick pointer of the texels.  A slot
// the server already gave to another node is a conflict.
__global__ void k_srv_adopt(uint64_t ntex, const uint32_t* __restrict__ tex, uint32_t ex,
                            uint32_t ey, uint32_t ez, uint32_t sx, uint32_t sy, uint32_t nslots,
                            uint32_t wide, uint32_t base, uint32_t* __restrict__ nodeSlot,
                            uint32_t* __restrict__ slotNode, uint32_t* __restrict__ conflict) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntex;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w0 = tex[2 * t], w1 = tex[2 * t + 1];
        if (!(w0 & kResidentBit)) continue;
        const uint32_t x = (((w1 >> 20) & 1023u) - 1u) / ex;
        const uint32_t y = (((w1 >> 10) & 1023u) - 1u) / ey;
        const uint32_t z = ((w1 & 1023u) - 1u) / ez;
        // a wide pool's word 1 is the slot itself
        const uint32_t s = wide ? w1 : (z * sy + y) * sx + x;
        if (s >= nslots) continue;
        if (atomicCAS(&slotNode[s], kNone, base + (uint32_t)t) != kNone) {
            atomicOr(conflict, 1u);
            continue;
        }
        nodeSlot[base + t] = s;
    }
}

__device__ __forceinline__ uint32_t parent_texel(uint32_t info, const TreeTab* T) {
    const uint32_t d = info >> 27;
    if (d == 0) return kNone;
    const uint32_t n = info & 0x07ffffffu;
    const uint32_t sh = d;  // 2^d nodes per axis
    const uint32_t m = (1u << sh) - 1u;
    const uint32_t x = n & m, y = (n >> sh) & m, z = n >> (2 * sh);
    const uint32_t p = (((z >> 1) << (2 * (sh - 1))) | ((y >> 1) << (sh - 1)) | (x >> 1));
    return T->texel[d - 1][p];
}

__device__ __forceinline__ bool non_const(uint32_t info, const TreeTab* T) {
    return T->types[info >> 27][info & 0x07ffffffu] == 1;
}

__device__ __forceinline__ void mark_pending(uint32_t g, unsigned long long emit,
                                             uint32_t* __restrict__ pend,
                                             unsigned long long* __restrict__ pendEmit) {
    if (atomicCAS(&pend[g], 0u, 2u) == 0u) pendEmit[g] = emit;
}

// A NON-CONST root without a brick becomes pending when its tree is
// attached (the pager queues a loaded tree's root, GigaVoxelsDatabasePager.cpp
// 300-310).
__global__ void k_srv_seed_root(const TreeTab* __restrict__ tabs, uint32_t tree, uint32_t base,
                                const uint32_t* __restrict__ nodeSlot, uint32_t* __restrict__ pend,
                                unsigned long long* __restrict__ pendEmit) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const TreeTab* T = tabs + tree;
    const uint32_t t = T->texel[0][0];
    if (t == kNone || T->types[0][0] != 1) return;
    if (nodeSlot[base + t] != kNone) return;
    mark_pending(base + t, (unsigned long long)wall_clock64(), pend, pendEmit);
}

// 1. touch used slots, mark new pending nodes (and their parents).
// tot: [0] uploads, [1] evictions, [2] request ids seen
__global__ void k_srv_touch(const uint32_t* __restrict__ ids, const uint64_t* __restrict__ times,
                            const uint32_t* __restrict__ count, uint32_t maxIds, uint32_t tree,
                            uint32_t base, uint32_t frame, const TreeTab* __restrict__ tabs,
                            const uint32_t* __restrict__ info, const uint32_t* __restrict__ nodeSlot,
                            uint32_t* __restrict__ slotLast, uint32_t* __restrict__ pend,
                            unsigned long long* __restrict__ pendEmit,
                            unsigned long long* __restrict__ tot) {
    const TreeTab* T = tabs + tree;
    if (!T->live) return;
    const uint32_t c = *count;
    const uint32_t nids = c < maxIds ? c : maxIds;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&tot[2], (unsigned long long)nids);
    const uint32_t ndx = T->ndim[0], ndy = T->ndim[1], ndz = T->ndim[2];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nids; k += gridDim.x * blockDim.x) {
        const uint32_t id = ids[k];
        const uint32_t x = ((id >> 20) & 1023u) - 1u, y = ((id >> 10) & 1023u) - 1u,
                       z = (id & 1023u) - 1u;
        if (x >= ndx || y >= ndy || z >= ndz) continue;
        const uint32_t g = base + (z * ndy + y) * ndx + x;
        const uint32_t inf = info[g];
        if (inf == kNone) continue;  // stale id outside the tree
        const uint32_t s = nodeSlot[g];
        if (s != kNone) {
            slotLast[s] = frame;
            continue;
        }
        if (!non_const(inf, T)) continue;
        const unsigned long long emit = (unsigned long long)times[k];
        mark_pending(g, emit, pend, pendEmit);
        const uint32_t p = parent_texel(inf, T);
        if (p == kNone) continue;
        const uint32_t gp = base + p;
        if (nodeSlot[gp] == kNone && non_const(info[gp], T)) mark_pending(gp, emit, pend, pendEmit);
    }
}

// Tree slots whose request lists a server frame reads, ascending: only their
// texels can have become pending in it, so the pending pass walks their
// ranges (n == 0: every texel).  Position v of the walk is texel
// s[v / cap] * cap + v % cap, ascending in v.
constexpr uint32_t kMaxPart = 64;
struct PartSlots {
    uint32_t n;
    uint32_t s[kMaxPart];
};
__device__ __forceinline__ uint64_t part_texel(const PartSlots& P, uint64_t cap, uint64_t v) {
    if (P.n == 0) return v;
    const uint64_t k = v / cap;
    return (uint64_t)P.s[k] * cap + (v - k * cap);
}

__global__ void k_srv_newflags(uint64_t nv, PartSlots P, uint64_t cap,
                               const uint32_t* __restrict__ pend, uint32_t* __restrict__ flag) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nv;
         v += (uint64_t)gridDim.x * blockDim.x)
        flag[v] = pend[part_texel(P, cap, v)] == 2u ? 1u : 0u;
}

// 2. append the new pending texels to the FIFO in ascending order.
// ctr: [0] head, [1] tail (monotonic), [2] admitted this frame
__global__ void k_srv_enqueue(uint64_t nv, PartSlots P, uint64_t tcap,
                              const uint32_t* __restrict__ flag,
                              const uint32_t* __restrict__ excl, uint32_t* __restrict__ pend,
                              const unsigned long long* __restrict__ pendEmit,
                              uint32_t* __restrict__ ring, unsigned long long* __restrict__ ringEmit,
                              uint32_t cap, const uint32_t* __restrict__ ctr) {
    const uint32_t tail = ctr[1];
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nv;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (!flag[v]) continue;
        const uint64_t t = part_texel(P, tcap, v);
        const uint32_t pos = (uint32_t)(((uint64_t)tail + excl[v]) % cap);
        ring[pos] = (uint32_t)t;
        ringEmit[pos] = pendEmit[t];
        pend[t] = 1u;
    }
}

__global__ void k_srv_tail(const uint32_t* __restrict__ excl, uint64_t ntex,
                           uint32_t* __restrict__ ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctr[1] += excl[ntex];
}

// 3a. slot keys and their histogram: lanes of a wave with one key add once
// to the workgroup's histogram in LDS, which adds its non-zero bins to the
// global one (most slots share a key: empty, or last used together).
__global__ void k_srv_keys(uint32_t nslots, uint32_t frame, const uint32_t* __restrict__ slotNode,
                           const uint32_t* __restrict__ slotLast, uint32_t* __restrict__ key,
                           uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kKeys];
    for (uint32_t i = threadIdx.x; i < kKeys; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63u);
    for (uint32_t base = blockIdx.x * blockDim.x; base < nslots; base += gridDim.x * blockDim.x) {
        const uint32_t s = base + threadIdx.x;
        const bool valid = s < nslots;
        uint32_t k = 0;
        if (valid) {
            if (slotNode[s] != kNone) {
                const uint32_t age = frame - slotLast[s];
                k = age == 0 ? 256u : 256u - (age < 255u ? age : 255u);
            }
            key[s] = k;
        }
        // the wave's most common keys (runs of equal ages) by ballot, at
        // most four rounds; the rest one LDS atomic per lane
        uint64_t pend = __ballot(valid);
        for (int round = 0; round < 4 && pend; ++round) {
            const int l = __ffsll((unsigned long long)pend) - 1;
            const uint32_t lk = __shfl(k, l);
            const uint64_t same = __ballot(valid && k == lk) & pend;
            if (lane == l) atomicAdd(&h[lk], (uint32_t)__popcll(same));
            pend &= ~same;
        }
        if ((pend >> lane) & 1ull) atomicAdd(&h[k], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kKeys; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
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
// req[i] = (tree << 8 | depth, brick, slot, texel).
__global__ void k_srv_admit(uint32_t nslots, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ excl, const uint32_t* __restrict__ sel,
                            uint32_t frame, const TreeTab* __restrict__ tabs, uint32_t cap,
                            const uint32_t* __restrict__ info, const uint32_t* __restrict__ ring,
                            const unsigned long long* __restrict__ ringEmit, uint32_t ringCap,
                            const uint32_t* __restrict__ ctr, uint32_t* __restrict__ nodeSlot,
                            uint32_t* __restrict__ slotNode, uint32_t* __restrict__ slotLast,
                            uint32_t* __restrict__ pend, uint4* __restrict__ req,
                            unsigned long long* __restrict__ reqEmit, uint32_t* __restrict__ nreq,
                            unsigned long long* __restrict__ tot) {
    const uint32_t n = sel[0];
    const uint32_t head = ctr[0];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x) {
        if (!flag[s]) continue;
        const uint32_t i = excl[s];
        if (i >= n) continue;
        const uint32_t pos = (uint32_t)(((uint64_t)head + i) % ringCap);
        const uint32_t g = ring[pos];
        const uint32_t old = slotNode[s];
        if (old != kNone) {
            nodeSlot[old] = kNone;
            const uint32_t to = old / cap;
            const TreeTab* To = tabs + to;
            if (To->live) To->tex[2 * (uint64_t)(old - to * cap)] &= ~kResidentBit;
            atomicAdd(&tot[1], 1ull);
        }
        slotNode[s] = g;
        slotLast[s] = frame;
        nodeSlot[g] = s;
        pend[g] = 0u;
        const uint32_t tr = g / cap;
        const uint32_t inf = info[g];
        const uint32_t d = inf >> 27, node = inf & 0x07ffffffu;
        req[i] = make_uint4((tr << 8) | d, tabs[tr].bidx[d][node], s, g - tr * cap);
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

// 5. copy each admitted brick from its tree's voxelizer result into its
// slot, set the node texel's resident bit and brick pointer, and record the
// device clock once the whole brick is written.
__global__ void k_srv_upload(const TreeTab* __restrict__ tabs, const uint4* __restrict__ req,
                             const uint32_t* __restrict__ nreq, uint32_t ex, uint32_t ey,
                             uint32_t ez, uint32_t sx, uint32_t sy, uint32_t dx, uint32_t dy,
                             uint32_t wide, uint32_t* __restrict__ col,
                             uint32_t* __restrict__ nrm, unsigned long long* __restrict__ done) {
    const uint32_t count = *nreq;
    const uint32_t nv = ex * ey * ez;
    for (uint32_t r = blockIdx.x; r < count; r += gridDim.x) {
        const uint4 q = req[r];
        const TreeTab* T = tabs + (q.x >> 8);
        const uint32_t d = q.x & 255u;
        const uint32_t* sc = T->col[d] + (uint64_t)q.y * nv;
        const uint32_t* sn = T->nrm[d] + (uint64_t)q.y * nv;
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
        if (threadIdx.x == 0) {
            T->tex[2 * (uint64_t)q.w] |= kResidentBit;
            T->tex[2 * (uint64_t)q.w + 1] = wide ? q.z : pack10(ox + 1, oy + 1, oz + 1);
        }
        __syncthreads();  // the whole brick is written
        if (threadIdx.x == 0) done[r] = (unsigned long long)wall_clock64();
    }
}

// Latency of each admitted request: copy done - emitted; counts the uploads.
__global__ void k_srv_latency(const uint32_t* __restrict__ nreq,
                              const unsigned long long* __restrict__ reqEmit,
                              const unsigned long long* __restrict__ done,
                              unsigned long long* __restrict__ log, uint64_t logCap,
                              const unsigned long long* __restrict__ logCount,
                              unsigned long long* __restrict__ tot) {
    const uint32_t n = *nreq;
    const unsigned long long base = *logCount;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&tot[0], (unsigned long long)n);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long k = base + i;
        if (k < logCap) log[k] = done[i] - reqEmit[i];
    }
}

__global__ void k_srv_logcount(const uint32_t* __restrict__ nreq,
                               unsigned long long* __restrict__ logCount) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *logCount += *nreq;
}

// Slots holding a brick, added into *count (one atomic per wave).
__global__ void k_srv_count_resident(uint32_t nslots, const uint32_t* __restrict__ slotNode,
                                     uint32_t* __restrict__ count) {
    uint32_t c = 0;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x)
        c += slotNode[s] != kNone ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
    if ((threadIdx.x & 63u) == 0u && c) atomicAdd(count, c);
}

// Detach: the tree's slots become empty (taken first) and its nodes lose
// their resident bits ...
__global__ void k_srv_detach_slots(uint32_t nslots, uint32_t lo, uint32_t hi,
                                   const TreeTab* __restrict__ T, uint32_t* __restrict__ slotNode,
                                   uint32_t* __restrict__ slotLast) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += gridDim.x * blockDim.x) {
        const uint32_t g = slotNode[s];
        if (g != kNone && g >= lo && g < hi) {
            T->tex[2 * (uint64_t)(g - lo)] &= ~kResidentBit;
            slotNode[s] = kNone;
            slotLast[s] = 0u;
        }
    }
}

// ... and its FIFO entries leave the queue, the others keeping their order:
// flag[i] for the i-th entry from the head that stays ...
__global__ void k_srv_ring_keep(const uint32_t* __restrict__ ring, uint32_t ringCap,
                                const uint32_t* __restrict__ ctr, uint32_t lo, uint32_t hi,
                                uint32_t* __restrict__ flag) {
    const uint32_t head = ctr[0], len = ctr[1] - ctr[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ringCap;
         i += gridDim.x * blockDim.x) {
        bool keep = false;
        if (i < len) {
            const uint32_t g = ring[(uint32_t)(((uint64_t)head + i) % ringCap)];
            keep = g < lo || g >= hi;
        }
        flag[i] = keep ? 1u : 0u;
    }
}

// ... moved to the front of the second ring, which then becomes the queue
// (head 0, tail = kept count).
__global__ void k_srv_ring_move(const uint32_t* __restrict__ ring,
                                const unsigned long long* __restrict__ ringEmit, uint32_t ringCap,
                                const uint32_t* __restrict__ ctr, const uint32_t* __restrict__ flag,
                                const uint32_t* __restrict__ excl, uint32_t* __restrict__ ring2,
                                unsigned long long* __restrict__ ringEmit2) {
    const uint32_t head = ctr[0];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ringCap;
         i += gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const uint32_t pos = (uint32_t)(((uint64_t)head + i) % ringCap);
        ring2[excl[i]] = ring[pos];
        ringEmit2[excl[i]] = ringEmit[pos];
    }
}

__global__ void k_srv_ring_reset(const uint32_t* __restrict__ excl, uint32_t ringCap,
                                 uint32_t* __restrict__ ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctr[0] = 0u;
        ctr[1] = excl[ringCap];
    }
}

}  // namespace

struct vox_server {
    vox_brick_pool* bp = nullptr;
    PoolRefs pool{};
    bool valid = false;
    uint32_t maxUp = 0, nslots = 0, frame = 0, maxTrees = 0;
    uint64_t cap = 0, ntex = 0;  // texels per tree, all trees
    hipStream_t st = nullptr;
    std::vector<vox_octree*> trees;
    DevBuf tabs, info, nodeSlot, slotNode, slotLast, pend, pendEmit, ring, ringEmit, ctr, tot, conflict;
    DevBuf ring2, ringEmit2;  // the queue after a detach's compaction
    DevBuf flagT, exclT, scanTmpT, key, hist, sel, flagS, eqExclS, exclS, scanTmpS;
    DevBuf req, reqEmit, nreq, done, log, logCount;
    DevBuf resident;          // slot count of the last stats call
    uint64_t logCap = 1u << 24;   // latencies kept (the first ones admitted)
    vox_server_stats stats{};
    // HIP-event pairs timing the server frames: the pairs of queued frames
    // (oldest first) and the free ones.  A frame settles the finished pairs
    // without waiting and takes a free pair, or makes one, so the host can
    // queue many frames ahead of the device (a pager queues a server frame
    // after every few ray casts); only past kMaxQueued pairs does it wait
    // for the oldest.
    struct EvPair { hipEvent_t a = nullptr, b = nullptr; };
    static constexpr size_t kMaxQueued = 512;
    std::deque<EvPair> evQueued;
    std::vector<EvPair> evFree;
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
    int slot_of(vox_octree* t) const {
        for (uint32_t i = 0; i < maxTrees; ++i)
            if (trees[i] == t) return (int)i;
        return -1;
    }
    // Add the times of queued frames, oldest first: all of them, waiting
    // for each, with all = true; else the finished ones, without waiting,
    // and then, while more than `keep` are queued, the oldest, waiting.
    void settle_timing(bool all = true, size_t keep = kMaxQueued) {
        while (!evQueued.empty()) {
            const EvPair pr = evQueued.front();
            if (all || evQueued.size() > keep) {
                (void)hipEventSynchronize(pr.b);
            } else if (hipEventQuery(pr.b) != hipSuccess) {
                (void)hipGetLastError();   // not ready is no error of the next launch
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, pr.a, pr.b);
            stats.ms_update_last = ms;
            stats.ms_update_total += ms;
            evQueued.pop_front();
            evFree.push_back(pr);
        }
    }
    // A pair for the next frame (settles what it can first).
    bool take_pair(EvPair& out) {
        settle_timing(false, kMaxQueued - 1);
        if (evFree.empty()) {
            EvPair pr;
            if (!check(hipEventCreate(&pr.a), "hipEventCreate")) return false;
            if (!check(hipEventCreate(&pr.b), "hipEventCreate")) {
                (void)hipEventDestroy(pr.a);
                return false;
            }
            evFree.push_back(pr);
        }
        out = evFree.back();
        evFree.pop_back();
        return true;
    }
    void destroy_events() {
        for (auto* v : {&evFree}) {
            for (auto& pr : *v) {
                (void)hipEventDestroy(pr.a);
                (void)hipEventDestroy(pr.b);
            }
            v->clear();
        }
        for (auto& pr : evQueued) {
            (void)hipEventDestroy(pr.a);
            (void)hipEventDestroy(pr.b);
        }
        evQueued.clear();
    }
};

namespace {

bool attach(vox_server* s, vox_octree* tree, uint32_t* id) {
    OctreeRefs o;
    if (!tree || !octree_refs(tree, o)) return s->fail("server: invalid octree");
    if (o.pool != s->bp) return s->fail("server: the octree uses another brick pool");
    if (o.levels > 9) return s->fail("server: more than 9 octree depths do not fit the texel table");
    const uint64_t ntex = (uint64_t)o.ndim[0] * o.ndim[1] * o.ndim[2];
    if (ntex > s->cap) return s->fail("server: the octree has more node texels than a tree slot");
    if (s->slot_of(tree) >= 0) return s->fail("server: the octree is attached already");
    const int i = s->slot_of(nullptr);
    if (i < 0) return s->fail("server: no free tree slot");
    hipStream_t st = s->st;
    if (!s->check(hipStreamWaitEvent(st, o.ready, 0), "wait for the octree")) return false;
    TreeTab v{};
    for (int d = 0; d < kMaxLevels; ++d) {
        v.types[d] = o.types[d];
        v.bidx[d] = o.bidx[d];
        v.texel[d] = o.texel[d];
        v.col[d] = o.src.col[d];
        v.nrm[d] = o.src.nrm[d];
    }
    v.tex = o.tex;
    for (int a = 0; a < 3; ++a) v.ndim[a] = o.ndim[a];
    v.live = 1;
    const uint32_t base = (uint32_t)(i * s->cap);
    launch_fill_u32(st, s->conflict.as<uint32_t>(), 1, 0u);
    k_srv_set_tab<<<1, 64, 0, st>>>(s->tabs.as<TreeTab>(), (uint32_t)i, v);
    for (int d = 0; d < o.levels; ++d) {
        const uint64_t n = 1ull << (3 * d);
        k_srv_info<<<grid_for(n), kB, 0, st>>>(n, (uint32_t)d, o.texel[d],
                                               s->info.as<uint32_t>() + base);
    }
    k_srv_adopt<<<grid_for(ntex), kB, 0, st>>>(
        ntex, o.tex, o.B[0] + 2, o.B[1] + 2, o.B[2] + 2, o.poolSlots[0], o.poolSlots[1], s->nslots,
        o.poolWide, base, s->nodeSlot.as<uint32_t>(), s->slotNode.as<uint32_t>(),
        s->conflict.as<uint32_t>());
    k_srv_seed_root<<<1, 64, 0, st>>>(s->tabs.as<TreeTab>(), (uint32_t)i, base,
                                      s->nodeSlot.as<uint32_t>(), s->pend.as<uint32_t>(),
                                      s->pendEmit.as<unsigned long long>());
    uint32_t conflict = 0;
    if (!s->check(hipGetLastError(), "attach kernels") ||
        !s->check(hipMemcpyAsync(&conflict, s->conflict.p, 4, hipMemcpyDeviceToHost, st), "flag") ||
        !s->check(hipStreamSynchronize(st), "synchronize"))
        return false;
    s->trees[i] = tree;
    if (conflict) {
        vox_server_detach(s, tree);
        return s->fail("server: a brick of the octree sits in a slot the server has given to "
                       "another node; build octrees for a shared pool with "
                       "VOX_TREE_NO_ROOT_BRICK");
    }
    if (id) *id = (uint32_t)i;
    return true;
}

}  // namespace

extern "C" {

vox_server* vox_server_create_shared(vox_brick_pool* bp, uint32_t max_trees,
                                     uint64_t max_texels_per_tree, uint32_t max_uploads_per_frame) {
    vox_server* s = new (std::nothrow) vox_server;
    if (!s) return nullptr;
    s->bp = bp;
    if (!bp || !pool_refs(bp, s->pool)) {
        s->fail("server: invalid brick pool");
        return s;
    }
    if (max_trees == 0 || max_texels_per_tree == 0 ||
        (uint64_t)max_trees * max_texels_per_tree >= 0xffffffffull) {
        s->fail("server: tree slots times texels per tree must be between 1 and 2^32 - 2");
        return s;
    }
    if (max_trees > (1u << 23)) {
        s->fail("server: too many tree slots");
        return s;
    }
    s->maxTrees = max_trees;
    s->cap = max_texels_per_tree;
    s->ntex = (uint64_t)max_trees * max_texels_per_tree;
    s->trees.assign(max_trees, nullptr);
    s->nslots = s->pool.slots[0] * s->pool.slots[1] * s->pool.slots[2];
    s->maxUp = max_uploads_per_frame ? max_uploads_per_frame : s->nslots;
    if (s->maxUp > s->nslots) s->maxUp = s->nslots;
    if (!s->check(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking), "hipStreamCreate"))
        return s;
    const uint64_t nt = s->ntex, ns = s->nslots;
    if (!s->tabs.ensure(sizeof(TreeTab) * max_trees, s->err, "tree table") ||
        !s->info.ensure(4 * nt, s->err, "info") || !s->nodeSlot.ensure(4 * nt, s->err, "nodeSlot") ||
        !s->slotNode.ensure(4 * ns, s->err, "slotNode") ||
        !s->slotLast.ensure(4 * ns, s->err, "slotLast") || !s->pend.ensure(4 * nt, s->err, "pend") ||
        !s->pendEmit.ensure(8 * nt, s->err, "pendEmit") || !s->ring.ensure(4 * nt, s->err, "ring") ||
        !s->ringEmit.ensure(8 * nt, s->err, "ringEmit") || !s->ctr.ensure(16, s->err, "ctr") ||
        !s->ring2.ensure(4 * nt, s->err, "ring") || !s->ringEmit2.ensure(8 * nt, s->err, "ringEmit") ||
        !s->tot.ensure(32, s->err, "totals") || !s->conflict.ensure(4, s->err, "conflict") ||
        !s->flagT.ensure(4 * nt, s->err, "flagT") || !s->exclT.ensure(4 * (nt + 1), s->err, "exclT") ||
        !ensure_scan_tmp(s->scanTmpT, nt, s->st, s->err, "scanTmpT") ||
        !s->key.ensure(4 * ns, s->err, "key") || !s->hist.ensure(4 * kKeys, s->err, "hist") ||
        !s->sel.ensure(16, s->err, "sel") || !s->flagS.ensure(4 * ns, s->err, "flagS") ||
        !s->eqExclS.ensure(4 * (ns + 1), s->err, "eqExclS") ||
        !s->exclS.ensure(4 * (ns + 1), s->err, "exclS") ||
        !ensure_scan_tmp(s->scanTmpS, ns, s->st, s->err, "scanTmpS") ||
        !s->req.ensure(16ull * s->maxUp, s->err, "req") ||
        !s->reqEmit.ensure(8ull * s->maxUp, s->err, "reqEmit") || !s->nreq.ensure(4, s->err, "nreq") ||
        !s->done.ensure(8ull * s->maxUp, s->err, "done") ||
        !s->log.ensure(8 * s->logCap, s->err, "latency log") ||
        !s->logCount.ensure(8, s->err, "latency count") ||
        !s->resident.ensure(4, s->err, "resident count"))
        return s;
    hipStream_t st = s->st;
    launch_fill_u32(st, s->tabs.as<uint32_t>(), sizeof(TreeTab) / 4 * max_trees, 0u);
    launch_fill_u32(st, s->info.as<uint32_t>(), nt, kNone);
    launch_fill_u32(st, s->nodeSlot.as<uint32_t>(), nt, kNone);
    launch_fill_u32(st, s->slotNode.as<uint32_t>(), ns, kNone);
    launch_fill_u32(st, s->slotLast.as<uint32_t>(), ns, 0u);
    launch_fill_u32(st, s->pend.as<uint32_t>(), nt, 0u);
    launch_fill_u32(st, s->ctr.as<uint32_t>(), 4, 0u);
    launch_fill_u32(st, s->tot.as<uint32_t>(), 8, 0u);
    launch_fill_u32(st, s->logCount.as<uint32_t>(), 2, 0u);
    // the pool's own stream may still be clearing it
    (void)hipStreamSynchronize(s->pool.st);
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

vox_server* vox_server_create(vox_octree* tree, uint32_t max_uploads_per_frame) {
    OctreeRefs o;
    if (!tree || !octree_refs(tree, o)) {
        vox_server* s = new (std::nothrow) vox_server;
        if (s) s->fail("server: invalid octree");
        return s;
    }
    const uint64_t ntex = (uint64_t)o.ndim[0] * o.ndim[1] * o.ndim[2];
    vox_server* s = vox_server_create_shared(o.pool, 1, ntex, max_uploads_per_frame);
    if (s && s->valid && !attach(s, tree, nullptr)) s->valid = false;
    return s;
}

void vox_server_destroy(vox_server* s) {
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    s->destroy_events();
    DevBuf* bufs[] = {&s->tabs, &s->info, &s->nodeSlot, &s->slotNode, &s->slotLast, &s->pend,
                      &s->pendEmit, &s->ring, &s->ringEmit, &s->ctr, &s->tot, &s->conflict,
                      &s->ring2, &s->ringEmit2,
                      &s->flagT, &s->exclT, &s->scanTmpT, &s->key, &s->hist, &s->sel, &s->flagS,
                      &s->eqExclS, &s->exclS, &s->scanTmpS, &s->req, &s->reqEmit, &s->nreq,
                      &s->done, &s->log, &s->logCount, &s->resident};
    for (DevBuf* b : bufs) b->release();
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

const char* vox_server_get_error(vox_server* s) {
    if (!s) return "null server";
    s->errCache = s->err.str();
    return s->errCache.c_str();
}

int vox_server_attach(vox_server* s, vox_octree* tree, uint32_t* id) {
    if (!s || !s->valid) return 0;
    return attach(s, tree, id) ? 1 : 0;
}

int vox_server_detach(vox_server* s, vox_octree* tree) {
    if (!s || !s->valid) return 0;
    const int i = s->slot_of(tree);
    if (!tree || i < 0) return s->fail("server: the octree is not attached"), 0;
    hipStream_t st = s->st;
    const uint32_t lo = (uint32_t)(i * s->cap), hi = (uint32_t)(lo + s->cap);
    k_srv_detach_slots<<<grid_for(s->nslots), kB, 0, st>>>(s->nslots, lo, hi,
                                                           s->tabs.as<TreeTab>() + i,
                                                           s->slotNode.as<uint32_t>(),
                                                           s->slotLast.as<uint32_t>());
    const uint32_t rc = (uint32_t)s->ntex;
    k_srv_ring_keep<<<grid_for(rc), kB, 0, st>>>(s->ring.as<uint32_t>(), rc, s->ctr.as<uint32_t>(),
                                                 lo, hi, s->flagT.as<uint32_t>());
    launch_scan(st, s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), rc, s->scanTmpT.as<uint32_t>());
    k_srv_ring_move<<<grid_for(rc), kB, 0, st>>>(
        s->ring.as<uint32_t>(), s->ringEmit.as<unsigned long long>(), rc, s->ctr.as<uint32_t>(),
        s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), s->ring2.as<uint32_t>(),
        s->ringEmit2.as<unsigned long long>());
    k_srv_ring_reset<<<1, 64, 0, st>>>(s->exclT.as<uint32_t>(), rc, s->ctr.as<uint32_t>());
    std::swap(s->ring, s->ring2);
    std::swap(s->ringEmit, s->ringEmit2);
    launch_fill_u32(st, s->nodeSlot.as<uint32_t>() + lo, s->cap, kNone);
    launch_fill_u32(st, s->pend.as<uint32_t>() + lo, s->cap, 0u);
    launch_fill_u32(st, s->info.as<uint32_t>() + lo, s->cap, kNone);
    k_srv_set_tab<<<1, 64, 0, st>>>(s->tabs.as<TreeTab>(), (uint32_t)i, TreeTab{});
    if (!octree_record_written(tree, st)) return s->fail("server: event record failed"), 0;
    s->trees[i] = nullptr;
    return s->check(hipGetLastError(), "detach kernels") &&
                   s->check(hipStreamSynchronize(st), "synchronize")
               ? 1
               : 0;
}

int vox_server_frame_multi(vox_server* s, vox_raycaster* const* rcs, uint32_t n) {
    if (!s || !s->valid) return 0;
    if (n && !rcs) return s->fail("server: no ray casters"), 0;
    std::vector<RequestRefs> lists(n);
    std::vector<uint32_t> slot(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (!raycaster_requests(rcs[k], lists[k])) return s->fail("server: invalid ray caster"), 0;
        const int i = s->slot_of(lists[k].tree);
        if (i < 0) return s->fail("server: a ray caster's octree is not attached"), 0;
        slot[k] = (uint32_t)i;
    }
    hipStream_t st = s->st;
    for (uint32_t k = 0; k < n; ++k)
        if (!s->check(hipStreamWaitEvent(st, lists[k].done, 0), "wait for a request list"))
            return 0;
    vox_server::EvPair evp;
    if (!s->take_pair(evp)) return 0;
    const uint32_t frame = ++s->frame;
    const uint64_t nt = s->ntex, ns = s->nslots;
    TreeTab* tabs = s->tabs.as<TreeTab>();
    (void)hipEventRecord(evp.a, st);
    for (uint32_t k = 0; k < n; ++k)
        k_srv_touch<<<grid_for(lists[k].maxIds), kB, 0, st>>>(
            lists[k].ids, lists[k].times, lists[k].count, lists[k].maxIds, slot[k],
            (uint32_t)(slot[k] * s->cap), frame, tabs, s->info.as<uint32_t>(),
            s->nodeSlot.as<uint32_t>(), s->slotLast.as<uint32_t>(), s->pend.as<uint32_t>(),
            s->pendEmit.as<unsigned long long>(), s->tot.as<unsigned long long>());
    // the pending pass over the texels of the trees in this frame's lists
    PartSlots ps{};
    {
        std::vector<uint32_t> u(slot.begin(), slot.end());
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if (!u.empty() && u.size() <= kMaxPart) {
            ps.n = (uint32_t)u.size();
            for (uint32_t k = 0; k < ps.n; ++k) ps.s[k] = u[k];
        }
    }
    const uint64_t nv = ps.n ? (uint64_t)ps.n * s->cap : nt;
    k_srv_newflags<<<grid_for(nv), kB, 0, st>>>(nv, ps, s->cap, s->pend.as<uint32_t>(),
                                                s->flagT.as<uint32_t>());
    launch_scan(st, s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), nv, s->scanTmpT.as<uint32_t>());
    k_srv_enqueue<<<grid_for(nv), kB, 0, st>>>(
        nv, ps, s->cap, s->flagT.as<uint32_t>(), s->exclT.as<uint32_t>(), s->pend.as<uint32_t>(),
        s->pendEmit.as<unsigned long long>(), s->ring.as<uint32_t>(),
        s->ringEmit.as<unsigned long long>(), (uint32_t)nt, s->ctr.as<uint32_t>());
    k_srv_tail<<<1, 64, 0, st>>>(s->exclT.as<uint32_t>(), nv, s->ctr.as<uint32_t>());
    launch_fill_u32(st, s->hist.as<uint32_t>(), kKeys, 0u);
    // a bounded grid: each workgroup adds its histogram to the global one
    // (257 atomics), so fewer workgroups mean fewer same-address atomics
    const unsigned kgrid = grid_for(ns) < 1024u ? grid_for(ns) : 1024u;
    k_srv_keys<<<kgrid, kB, 0, st>>>((uint32_t)ns, frame, s->slotNode.as<uint32_t>(),
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
        frame, tabs, (uint32_t)s->cap, s->info.as<uint32_t>(), s->ring.as<uint32_t>(),
        s->ringEmit.as<unsigned long long>(), (uint32_t)nt, s->ctr.as<uint32_t>(),
        s->nodeSlot.as<uint32_t>(), s->slotNode.as<uint32_t>(), s->slotLast.as<uint32_t>(),
        s->pend.as<uint32_t>(), s->req.as<uint4>(), s->reqEmit.as<unsigned long long>(),
        s->nreq.as<uint32_t>(), s->tot.as<unsigned long long>());
    k_srv_advance<<<1, 64, 0, st>>>(s->sel.as<uint32_t>(), s->ctr.as<uint32_t>());
    const PoolRefs& P = s->pool;
    const unsigned ugrid = s->maxUp < 4096u ? s->maxUp : 4096u;
    k_srv_upload<<<ugrid, kB, 0, st>>>(tabs, s->req.as<uint4>(), s->nreq.as<uint32_t>(),
                                       P.B[0] + 2, P.B[1] + 2, P.B[2] + 2, P.slots[0], P.slots[1],
                                       P.dims[0], P.dims[1], P.wide, P.col, P.nrm,
                                       s->done.as<unsigned long long>());
    k_srv_latency<<<grid_for(s->maxUp), kB, 0, st>>>(
        s->nreq.as<uint32_t>(), s->reqEmit.as<unsigned long long>(),
        s->done.as<unsigned long long>(), s->log.as<unsigned long long>(), s->logCap,
        s->logCount.as<unsigned long long>(), s->tot.as<unsigned long long>());
    k_srv_logcount<<<1, 64, 0, st>>>(s->nreq.as<uint32_t>(), s->logCount.as<unsigned long long>());
    (void)hipEventRecord(evp.b, st);
    s->evQueued.push_back(evp);
    if (!s->check(hipGetLastError(), "server kernels")) return 0;
    // ray casts of the next frame wait for the node textures and pool
    for (vox_octree* t : s->trees)
        if (t && !octree_record_written(t, st)) return s->fail("server: event record failed"), 0;
    s->stats.frames = frame;
    return 1;
}

int vox_server_frame(vox_server* s, vox_raycaster* rc) {
    return vox_server_frame_multi(s, &rc, 1);
}

int vox_server_get_stats(vox_server* s, vox_server_stats* out) {
    if (!s || !s->valid || !out) return 0;
    uint32_t ctr[4] = {0, 0, 0, 0};
    unsigned long long tot[4] = {0, 0, 0, 0};
    uint32_t resident = 0;
    launch_fill_u32(s->st, s->resident.as<uint32_t>(), 1, 0u);
    k_srv_count_resident<<<grid_for(s->nslots < 4096ull * kB ? s->nslots : 4096ull * kB), kB, 0,
                           s->st>>>((uint32_t)s->nslots, s->slotNode.as<uint32_t>(),
                                    s->resident.as<uint32_t>());
    if (!s->check(hipGetLastError(), "resident count") ||
        !s->check(hipMemcpyAsync(ctr, s->ctr.p, 16, hipMemcpyDeviceToHost, s->st), "counters") ||
        !s->check(hipMemcpyAsync(tot, s->tot.p, 32, hipMemcpyDeviceToHost, s->st), "totals") ||
        !s->check(hipMemcpyAsync(&resident, s->resident.p, 4, hipMemcpyDeviceToHost, s->st),
                  "resident count") ||
        !s->check(hipStreamSynchronize(s->st), "synchronize"))
        return 0;
    s->settle_timing();
    s->stats.uploads = tot[0];
    s->stats.evictions = tot[1];
    s->stats.requests = tot[2];
    s->stats.pending = ctr[1] - ctr[0];
    s->stats.last_uploads = ctr[2];
    s->stats.resident = resident;
    *out = s->stats;
    return 1;
}

int vox_server_latencies(vox_server* s, uint64_t* out, uint64_t cap, uint64_t* n) {
    if (!s || !s->valid || !n) return 0;
    unsigned long long cnt = 0;
    if (!s->check(hipMemcpyAsync(&cnt, s->logCount.p, 8, hipMemcpyDeviceToHost, s->st), "count") ||
        !s->check(hipStreamSynchronize(s->st), "synchronize"))
        return 0;
    const uint64_t kept = cnt < s->logCap ? cnt : s->logCap;
    *n = kept;
    uint64_t m = kept;
    if (m > cap) m = cap;
    if (m && out && !s->check(hipMemcpy(out, s->log.p, 8 * m, hipMemcpyDeviceToHost), "log"))
        return 0;
    return 1;
}

int vox_server_state(vox_server* s, const uint32_t** node_slot, const uint32_t** slot_node,
                     uint64_t* ntexels, uint32_t* nslots) {
    if (!s || !s->valid) return 0;
    if (!s->check(hipStreamSynchronize(s->st), "synchronize")) return 0;
    if (node_slot) *node_slot = s->nodeSlot.as<uint32_t>();
    if (slot_node) *slot_node = s->slotNode.as<uint32_t>();
    if (ntexels) *ntexels = s->ntex;
    if (nslots) *nslots = s->nslots;
    return 1;
}

}  // extern "C"
