/*
 * voxmi_pool.h -- C ABI of the MI355X GigaVoxels node pool and brick pool:
 * the runtime-side layout a voxelized octree is served from.
 *
 * The reference builds these on the host from the voxelizer's files:
 * GigaVoxelsReader::Load parses tree.gvx depth-first and appends one block of
 * 8 nodes per node that has children (VoxViz/GigaVoxels/GigaVoxelsReader.cpp
 * 277-514, GigaVoxelsOctTreeNodePool.cpp 49-74), BrickPool::initBricks
 * uploads the root's brick (GigaVoxelsBrickPool.cpp 270-345) and
 * OctTreeNodePool::create3DTexture lays the blocks out in an RG32UI 3D texture
 * (GigaVoxelsOctTreeNodePool.cpp 90-223).  Here the same texture is built on
 * the device directly from the voxelizer's in-HBM result, with the same
 * block order (depth-first allocation), the same texture dimensions and the
 * same node bit encoding (GigaVoxelsOctTree.cpp 1309-1383):
 *
 *   word 0: bit 31 brick resident, bit 30 constant, bits 29-20 / 19-10 / 9-0
 *           child block texel x / y / z (0 = no children)
 *   word 1: brick texel x / y / z of a resident brick (slot origin + 1), or
 *           the constant RGBA8 colour (R in the low byte) of a constant node
 *           (a wide pool, vox_brick_pool_create_wide: the slot index)
 *
 * The brick pool is the reference's 3D texture of (B+2)*slots texels per
 * axis, x fastest, one RGBA8 colour texel and one RGB8 normal texel (in a
 * 32-bit word) per voxel; slots are numbered x fastest (initSpecial,
 * GigaVoxelsBrickPool.cpp 153-250).
 */
#ifndef VOXMI_POOL_H
#define VOXMI_POOL_H

#include <stdint.h>

#include "voxmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_brick_pool vox_brick_pool;
typedef struct vox_octree vox_octree;

/* Word 1 of constant nodes.  The reference's reader parses the colour from
 * the tree.gvx text (c/255 printed with 6 significant digits), computes
 * uchar(value * 255.0f) and never stores it (Node::setConstantValue,
 * GigaVoxelsOctTree.cpp 1611-1621), so its node pool holds 0 there.
 *   default                   the colour the reader computes from the text
 *                             (122 of the 256 channel values come back one
 *                             lower through the text), i.e. the reader with
 *                             the store added;
 *   VOX_TREE_REFERENCE_CONST  0, as the reference's node pool holds;
 *   VOX_TREE_EXACT_CONST      the voxelizer's colour itself. */
#define VOX_TREE_REFERENCE_CONST 1u
#define VOX_TREE_EXACT_CONST 2u
/* The root's brick is not uploaded at creation; a brick server the tree is
 * attached to admits it through its LRU like any other brick, as the
 * reference's pager does for the trees it loads (GigaVoxelsDatabasePager.cpp
 * 300-310). */
#define VOX_TREE_NO_ROOT_BRICK 4u

typedef struct vox_brick_pool_view {
    uint32_t  brick_dim[3];     /* B (stored bricks are B+2 per axis)         */
    uint32_t  slots[3];         /* slots per axis                             */
    uint32_t  dims[3];          /* texels per axis = (B+2) * slots            */
    uint32_t* colors;           /* device, dims product RGBA8 texels          */
    uint32_t* normals;          /* device, dims product RGB8 texels           */
    uint32_t  resident;         /* slots holding a brick                      */
    uint32_t  wide;             /* 1: vox_brick_pool_create_wide layout       */
} vox_brick_pool_view;

typedef struct vox_octree_view {
    int32_t   levels;
    uint32_t  node_dim[3];      /* node texture texels per axis               */
    uint64_t  node_blocks;      /* blocks of 8 nodes (block 0 holds the root) */
    uint32_t* nodes;            /* device, 2 words per texel, x fastest       */
    /* per depth d: texel index of each of the 8^d nodes (x fastest), or
     * 0xffffffff for nodes below a LEAF-CONST node (not in the tree) */
    const uint32_t* node_texel[VOX_MAX_LEVELS];
} vox_octree_view;

/* BrickPool::initSpecial (GigaVoxelsBrickPool.cpp 153-250).  slots may be
 * NULL: the renderer's choice, 8^3 for bricks wider than 14, else 16^3
 * (GigaVoxelsRenderer.cpp 259-272).  Texel coordinates are 10-bit fields,
 * so (B+2)*slots must stay <= 1024 per axis. */
vox_brick_pool* vox_brick_pool_create(const uint32_t brick_dim[3], const uint32_t slots[3]);
/* A brick pool beyond the 10-bit texel pointers of the reference's node
 * format: nslots slots (up to 2^30) stacked along z, slot s holding texels
 * [s * (B+2)^3, (s + 1) * (B+2)^3), x fastest.  Word 1 of a node whose brick
 * is resident holds the slot index s instead of texel coordinates, and the
 * ray cast samples each brick in its own (B+2)^3 texel frame (normalised
 * coordinates of the brick, the same trilinear filter).  An extension of the
 * reference (whose pool is one 3D texture, at most 1024 texels per axis with
 * 10-bit pointers, i.e. 8.2 GB of 18^3 colour + normal bricks): for pools of
 * 16 GB and more.  Nodes built on it are not readable by the reference's
 * shader. */
vox_brick_pool* vox_brick_pool_create_wide(const uint32_t brick_dim[3], uint64_t nslots);
/* A wide pool of RGBA8 colour bricks only (vox_brick_pool_view.normals is
 * NULL; uploads skip the normals): twice the slots in the same memory, for
 * a ray cast that samples colour only, as this library's does.  The paged
 * workloads' pools are sized this way (16 GB of 18^3 RGBA8 bricks hold
 * 686 K slots). */
vox_brick_pool* vox_brick_pool_create_colour(const uint32_t brick_dim[3], uint64_t nslots);
void vox_brick_pool_destroy(vox_brick_pool* bp);
const char* vox_brick_pool_get_error(vox_brick_pool* bp);
int vox_brick_pool_get_view(vox_brick_pool* bp, vox_brick_pool_view* out);

/* GigaVoxelsReader::Load of one octree + GigaVoxelsOctTree::
 * uploadInitialBricks + initNodePoolAndNodeUsageListProcessor
 * (GigaVoxelsRenderer.cpp 285-286): builds the node texture from the device
 * result of vox_generate and uploads the root's brick into the pool when the
 * root is NON-CONST.  The tree's arrays are copied; its bricks are read from
 * the result when they are uploaded, so the voxelizer handle must not
 * generate again while the octree is served.  Returns NULL only when out of
 * host memory; errors are reported by vox_octree_get_error. */
vox_octree* vox_octree_create(vox_brick_pool* bp, const vox_result* tree, uint32_t flags);
void vox_octree_destroy(vox_octree* t);
const char* vox_octree_get_error(vox_octree* t);
int vox_octree_get_view(vox_octree* t, vox_octree_view* out);

/* Copy n bytes from a device pointer of a view to host memory, after all
 * device work queued so far (every stream) has finished. */
int vox_pool_copy_to_host(void* dst, const void* src_device, uint64_t nbytes);

#ifdef __cplusplus
}
#endif

#endif
