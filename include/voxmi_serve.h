/*
 * voxmi_serve.h -- C ABI of the sparse-octree ray cast that produces brick
 * requests, and of the device-resident brick server that admits them into
 * the brick pool with LRU eviction.
 *
 * Ray cast: the front-to-back traversal of VoxViz/Shaders/GigaVoxels.frag
 * (main 544-805, TraverseConstant 219-266, TraverseBrick 311-375,
 * AddToNodeUsageList 508-542) as a HIP kernel, one thread per pixel, over the
 * node texture and brick pool of voxmi_pool.h.  Each pixel records up to 12
 * node ids ((x+1)<<20 | (y+1)<<10 | (z+1) of the node texel, 0 = none,
 * GigaVoxels.frag 508-525) in the 2x2-pixel interleave of the shader
 * (571-577).  The ids are then deduplicated on the device (wave-level first,
 * then one flag per node texel) into the frame's request list, replacing the
 * selection-mask / histopyramid / compaction passes and the host-side
 * std::map of the reference (GigaVoxelsRenderer.cpp 2048-2110,
 * GigaVoxelsOctTree.cpp 430-652).
 *
 * Brick server: per frame, the nodes of the request list that hold a brick
 * in the pool are marked used; NON-CONST nodes without a resident brick are
 * admitted into the least recently used slots (BrickPool::update,
 * GigaVoxelsBrickPool.cpp 561-682), their bricks copied from the voxelizer's
 * result in HBM and their node texels updated.  Requests that find no
 * evictable slot stay queued and are admitted in a later frame.
 *
 * Ordering.  Every ray caster and server runs on its own HIP stream and
 * returns without waiting for the device.  A ray-cast frame starts after the
 * last write of its octree's node texture and brick pool (the octree build or
 * a server frame); a server frame starts after the request lists it is given.
 * Ray casters of one frame therefore overlap on the device with each other
 * and with unrelated work, and a host loop of "ray cast every stream, then
 * one server frame" runs without host round trips.  All ray casts reading a
 * pool must be given to the server frame that follows them (or have
 * finished): the server overwrites slots.  Reading results
 * (vox_raycaster_get_frame, vox_server_get_stats, vox_server_state,
 * vox_pool_copy_to_host) waits for the work they read.
 */
#ifndef VOXMI_SERVE_H
#define VOXMI_SERVE_H

#include <stdint.h>

#include "voxmi_pool.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_raycaster vox_raycaster;

/* A pinhole camera in world units.  look and up are unit vectors and
 * orthogonal; left = up x look (the reference's Camera::getLeft). */
typedef struct vox_camera {
    float    position[3];
    float    look[3];
    float    up[3];
    float    fov_y_deg;         /* vertical field of view                    */
    int32_t  width, height;     /* viewport in pixels                        */
    uint32_t frame;             /* frame counter (Camera::getFrameCount)     */
} vox_camera;

typedef struct vox_raycast_params {
    float    vol_min[3];        /* world box of the octree                   */
    float    vol_max[3];
    float    ray_step;          /* RayStepSize in unit-volume coordinates;
                                   0: half a leaf voxel, as the renderer sets
                                   it (GigaVoxelsRenderer.cpp 1705)          */
    float    lod_scalar;        /* LodScalar; 0: 10 (GigaVoxelsRenderer.cpp 151) */
    uint32_t jitter_seed;       /* 32x32 jitter table (JitterTexSize = 32)   */
    uint32_t max_traversals;    /* bound on root restarts per ray; 0: 4096.
                                   Not in the reference; every ray ends.     */
    int32_t  rect[4];           /* pixels x0, y0, x1, y1 (exclusive) cast this
                                   frame; all zero: the whole viewport.  The
                                   reference shades only the pixels the
                                   tree's box covers; a rectangle holding the
                                   box's projection gives the same result
                                   there.  Pixels outside are not written.   */
} vox_raycast_params;

typedef struct vox_frame_view {
    int32_t   width, height;
    uint32_t* rgba;             /* device, width*height RGBA8 (R low byte)   */
    uint32_t* usage;            /* device, width*height*12 node ids          */
    uint32_t* requests;         /* device, the frame's distinct node ids     */
    uint32_t  nrequests;        /* capped at max_requests                    */
    uint32_t  nrequests_total;  /* distinct ids before the cap               */
    double    ms_raycast;       /* HIP-event time of the ray-cast kernel     */
    double    ms_requests;      /* HIP-event time of the deduplication       */
    const float* rgba_f32;      /* device, width*height*4 floats: FragColor  */
} vox_frame_view;

/* Ray caster over one octree (its node texture and brick pool).  The frame
 * buffers are sized for max_width x max_height; max_requests caps the
 * request list (the reference caps the compacted list at 32768,
 * GigaVoxelsRenderer.cpp 2063-2064). */
vox_raycaster* vox_raycaster_create(vox_octree* tree, int32_t max_width, int32_t max_height,
                                    uint32_t max_requests);
void vox_raycaster_destroy(vox_raycaster* rc);
const char* vox_raycaster_get_error(vox_raycaster* rc);
/* One frame: ray cast, then request deduplication, queued on the ray
 * caster's stream. */
int vox_raycast_frame(vox_raycaster* rc, const vox_camera* cam, const vox_raycast_params* prm);
/* The last frame's buffers, counts and timings (waits for the frame). */
int vox_raycaster_get_frame(vox_raycaster* rc, vox_frame_view* out);
/* One front-to-back frame over the octrees of n ray casters (one camera,
 * per octree its own parameters: box, pixel rectangle; one jitter seed),
 * queued on the first caster's stream.  rcs must list the octrees in an
 * order every ray meets them in (boxes that tile space, sorted outward from
 * the camera).  Each pixel casts its ray through each octree whose rectangle
 * holds it exactly as vox_raycast_frame does, and the casts' FragColors are
 * accumulated as vox_composite's far-to-near blend would accumulate them
 * (out = sum of T * src * src.a per channel, T the product of (1 - src.a) of
 * the octrees in front), into frame: width*height RGBA32F pixels on the
 * device, every pixel written.  A pixel whose T falls to t_stop or below
 * (t_stop 0: never, even at T = 0) casts no further octree, so their bricks are not requested
 * through it; with t_stop = 0 the frame equals vox_composite's up to fp32
 * rounding.  Each caster gets the request list of its octree's casts (read
 * by vox_server_frame / _multi after this frame); its colour and usage
 * buffers are not written.  A frame may be cast in several passes over
 * consecutive runs of that order (a server frame between them serves the
 * nearer octrees' requests first): transmittance (width*height floats on the
 * device, or NULL) receives each pixel's T, and resume = 1 continues from
 * frame and transmittance as the previous pass left them; a pass writes only
 * the union of its octrees' pixel rectangles, and resume = 0 first clears
 * the whole frame (and sets transmittance to 1).  Whatever wrote frame last
 * (the previous pass, or the last frame) must be ordered before the call on
 * the device, as a server frame between them orders it.  Not in the reference, which rasterizes and casts every
 * paged octree separately (GigaVoxelsRenderer.cpp 1019-1040). */
int vox_raycast_frame_multi(vox_raycaster* const* rcs, uint32_t n, const vox_camera* cam,
                            const vox_raycast_params* prms, float* frame, float* transmittance,
                            float t_stop, int32_t resume);
/* Blend the last frames of n ray casters, in the order given (far to near,
 * as the renderer draws its list), into dst: width*height RGBA32F pixels on
 * the device, the renderer's colour target (GigaVoxelsRenderer.cpp 511-515).
 * Each caster's pixel rectangle is blended with glBlendFunc(GL_SRC_ALPHA,
 * GL_ONE_MINUS_SRC_ALPHA) (GigaVoxelsRenderer.cpp 1019-1022): per channel,
 * alpha included, out = src * src.a + dst * (1 - src.a) in fp32, src being
 * the caster's FragColor.  clear = 1 first zeroes dst (glClearBufferfv).
 * Queued after the casters' frames; their next frames wait for it. */
int vox_composite(vox_raycaster* const* rcs, uint32_t n, float* dst, int32_t width,
                  int32_t height, int32_t clear);

/* ---- brick server ---------------------------------------------------------
 * Per frame (vox_server_frame), on the device:
 *  1. every requested node with a resident brick marks its slot used this
 *     frame (BrickPool::notifyUsed, GigaVoxelsBrickPool.cpp 406-427); every
 *     requested NON-CONST node without one becomes pending, and so does its
 *     parent when that is NON-CONST without a brick (BrickPool::update uploads
 *     the parent too, GigaVoxelsBrickPool.cpp 631-676);
 *  2. the frame's new pending nodes join a FIFO in ascending texel order;
 *  3. up to max_uploads_per_frame FIFO entries are admitted into slots not
 *     used this frame: empty slots first, then least recently used, ties by
 *     slot index; an evicted node loses its resident bit;
 *  4. the admitted bricks are copied from the voxelizer result into their
 *     slots and their node texels get the resident bit and brick pointer.
 * Entries that find no slot stay in the FIFO for later frames.  The latency
 * of an admitted request is the device clock from the ray cast appending it
 * to the end of its brick copy. */
typedef struct vox_server vox_server;

typedef struct vox_server_stats {
    uint64_t frames;
    uint64_t requests;          /* distinct request ids seen, over frames    */
    uint64_t uploads;           /* bricks copied into the pool               */
    uint64_t evictions;
    uint32_t pending;           /* FIFO entries waiting for a slot           */
    uint32_t resident;          /* slots holding a brick                     */
    uint32_t last_uploads;      /* admitted in the last frame                */
    double   ms_update_last;    /* HIP-event time of the last update         */
    double   ms_update_total;
    double   clock_khz;         /* device clock the latencies are counted in */
} vox_server_stats;

/* Server over one octree and its brick pool, adopting the root brick the
 * octree build uploaded.  max_uploads_per_frame bounds the admissions per
 * frame (0: the pool's slot count). */
vox_server* vox_server_create(vox_octree* tree, uint32_t max_uploads_per_frame);
/* Server over a brick pool shared by up to max_trees octrees of at most
 * max_texels_per_tree node texels each (the reference's one BrickPool with
 * one LRU list for every tree the pager has loaded, GigaVoxelsBrickPool.cpp
 * 406-449).  Nodes are named by global texel: tree slot * max_texels_per_tree
 * + texel (vox_server_state, FIFO order).  Not in the reference as an
 * object of its own. */
vox_server* vox_server_create_shared(vox_brick_pool* bp, uint32_t max_trees,
                                     uint64_t max_texels_per_tree, uint32_t max_uploads_per_frame);
/* Attach an octree built on the server's pool; *id receives its tree slot.
 * A NON-CONST root without a resident brick becomes pending at once, as the
 * pager queues the root of every tree it loads (GigaVoxelsDatabasePager.cpp
 * 300-310); build such octrees with VOX_TREE_NO_ROOT_BRICK.  Bricks already
 * resident in the octree's texture are adopted (an error if their slot is
 * taken).  Waits for the device. */
int vox_server_attach(vox_server* s, vox_octree* tree, uint32_t* id);
/* Detach an octree (the pager unloading a tree): its slots become empty and
 * are taken first, as BrickPool::notifyDeleted moves them to the front of
 * the LRU list (GigaVoxelsBrickPool.cpp 426-449), and its nodes lose their
 * resident bits; its queued requests are voided (an admission that meets
 * one leaves the slot as it was).  Waits for the device; the octree may then
 * be destroyed (its ray casters first) or attached again. */
int vox_server_detach(vox_server* s, vox_octree* tree);
void vox_server_destroy(vox_server* s);
const char* vox_server_get_error(vox_server* s);
/* Serve the request list of the ray caster's last frame (queued). */
int vox_server_frame(vox_server* s, vox_raycaster* rc);
/* Serve the request lists of n ray casters' last frames -- camera streams,
 * or the trees of one view -- in one admission step: the lists are touched
 * in the order given, then new pending nodes join the FIFO in ascending
 * global texel order.  Each ray caster's octree must be attached. */
int vox_server_frame_multi(vox_server* s, vox_raycaster* const* rcs, uint32_t n);
int vox_server_get_stats(vox_server* s, vox_server_stats* out);
/* Latencies (device clock ticks, see clock_khz) of the admitted requests
 * since creation, in admission order: the first 2^24 are kept; *n = the
 * number kept, min(*n, cap) of them are copied (out may be NULL to ask for
 * *n).  vox_server_get_stats counts every upload. */
int vox_server_latencies(vox_server* s, uint64_t* out, uint64_t cap, uint64_t* n);
/* Slot of each global node texel (0xffffffff: none) and global node texel
 * of each slot, device pointers (waits for the server's queued work). */
int vox_server_state(vox_server* s, const uint32_t** node_slot, const uint32_t** slot_node,
                     uint64_t* ntexels, uint32_t* nslots);

#ifdef __cplusplus
}
#endif

#endif
