/*
 * ref_pool.h -- CPU restatement of how the GigaVoxels runtime turns the
 * voxelizer's files into its GPU node pool and brick pool
 * (apecoraro/VoxVizThesis, VoxViz/GigaVoxels).
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle the HIP node-pool builder is checked
 * against; only tests/ and __graft_entry__.smoke() load it.
 *
 * Follows, in the reference's own data flow:
 *   GigaVoxelsReader.cpp 374-514 (LoadGVX) and 277-372 (LoadOctTreeNode):
 *     recursive depth-first parse of tree.gvx; a child block of 8 nodes is
 *     appended to the node vector for every node that is neither at the last
 *     depth nor LEAF-CONST, before its children are parsed;
 *   GigaVoxelsReader.cpp 106-221 (LoadBinaryBricks): voxels_<d>.gvb, bricks
 *     in Brick= index order, EOF required after the last brick;
 *   GigaVoxelsOctTreeNodePool.cpp 49-74, 90-223 (allocateChildNodeBlock,
 *     create3DTexture): texture side floor(pow(n, 1/3.0)) forced even, x grown
 *     by 2 until it holds n nodes; block b at 2*(b % bx, (b / bx) % by,
 *     b / (bx*by)); nodes of a block x fastest;
 *   GigaVoxelsOctTree.cpp 1309-1383, 1611-1621 (Node bit encoding);
 *   GigaVoxelsBrickPool.cpp 153-250, 270-345, 729-790 (initSpecial,
 *     initBricks(root, 1) as called by GigaVoxelsOctTree.cpp 1703-1706):
 *     the root's brick, when the root is NON-CONST, goes to slot 0 before the
 *     node texture is built.
 */
#ifndef VOXMI_REF_POOL_H
#define VOXMI_REF_POOL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_pool ref_pool;

/* flags bit 0: store the constant colour of CONST / LEAF-CONST nodes in the
 * second word.  The reader computes it and drops it (Node::setConstantValue,
 * GigaVoxelsOctTree.cpp 1611-1621), so the reference's node pool holds 0
 * there; with the bit set the oracle stores uchar(float(text) * 255.0f). */
#define REF_POOL_STORE_CONST 1

/* Load dir/tree.gvx and dir/voxels_<d>.gvb.  slots = brick-pool slots per
 * axis (the renderer uses 8^3 for bricks wider than 14, else 16^3).
 * Returns NULL with a message in ref_pool_error() on a parse error. */
ref_pool* ref_pool_load(const char* dir, const uint32_t slots[3], int32_t flags);
const char* ref_pool_error(void);

void            ref_pool_node_dims(const ref_pool* p, uint32_t dims[3], uint64_t* blocks);
const uint32_t* ref_pool_nodes(const ref_pool* p);        /* 2 u32 per texel, x fastest */
void            ref_pool_brick_dims(const ref_pool* p, uint32_t dims[3]);
const uint32_t* ref_pool_brick_colors(const ref_pool* p);  /* RGBA8 texels, x fastest */
const uint32_t* ref_pool_brick_normals(const ref_pool* p); /* RGB8 texels in u32 */
int32_t         ref_pool_depths(const ref_pool* p);
uint64_t        ref_pool_file_bricks(const ref_pool* p, int32_t depth);
const uint32_t* ref_pool_file_brick_colors(const ref_pool* p, int32_t depth);
const uint32_t* ref_pool_file_brick_normals(const ref_pool* p, int32_t depth);
/* Compressed .gvb files (flag 1): the payloads of depth d as read, colours
 * (normals = 0) or normals (normals = 1); the root's brick is not uploaded
 * (the pool here holds uncompressed texels). */
int32_t         ref_pool_compressed(const ref_pool* p);
const uint8_t*  ref_pool_file_compressed(const ref_pool* p, int32_t depth, int32_t normals,
                                         uint64_t* nbytes);
void            ref_pool_free(ref_pool* p);

#ifdef __cplusplus
}
#endif

#endif
