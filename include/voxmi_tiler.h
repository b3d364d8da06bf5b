/*
 * voxmi_tiler.h -- assignment of triangles to the grid of octree tiles, the
 * device part of PagedGigaVoxelOctTreeGenerator (main.cu 3811-4507).
 *
 * The reference splits the input's x/y extent into tiles of
 * maxVoxelMipMapSize metres (one octree each, grid_x_y_z) and assigns every
 * triangle to the tiles its vertices reach, a vertex within the voxel border
 * of a tile edge reaching the neighbouring tile too
 * (TriangleOctTreeGrid::assignTriangles / computeMinMaxGridCell, main.cu
 * 3986-4107).  vox_tiler_assign does this on the device for a triangle soup
 * in HBM and returns, per tile, the indices of its triangles and the z extent
 * of their vertices.  The per-tile voxelization and root.gvp are driven by
 * the host (voxvizthesis_amd/tiler.py).
 */
#ifndef VOXMI_TILER_H
#define VOXMI_TILER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_tiler vox_tiler;

typedef struct vox_tile_grid {
    float    origin[2];         /* x, y of the input's bounding-box minimum    */
    float    size[2];           /* tile edge in metres (voxel size * N)        */
    float    border[2];         /* voxel border in metres (voxel size * extra) */
    uint32_t dim[2];            /* tiles along x and y                         */
} vox_tile_grid;

typedef struct vox_tile_lists {
    uint64_t        ntiles;     /* dim[0] * dim[1], tile index y * dim[0] + x  */
    uint64_t        pairs;      /* (tile, triangle) entries                    */
    const uint32_t* tile_start; /* device, ntiles + 1 offsets                  */
    const uint32_t* tile_tris;  /* device, triangle indices of each tile       */
    const float*    tile_zmin;  /* device, per tile: min / max vertex z of its */
    const float*    tile_zmax;  /*   triangles (NaN for an empty tile)         */
} vox_tile_lists;

vox_tiler* vox_tiler_create(void);
void vox_tiler_destroy(vox_tiler* t);
const char* vox_tiler_get_error(vox_tiler* t);
/* verts: 9 floats per triangle, a device pointer (device = 1) or host
 * memory copied in first (device = 0).  Triangle indices within a tile come
 * out in ascending (input) order, the order the reference appends them in
 * (the voxelizer's per-voxel triangle slots depend on it). */
int vox_tiler_assign(vox_tiler* t, const float* verts, uint64_t ntri, int32_t device,
                     const vox_tile_grid* grid);
int vox_tiler_get(vox_tiler* t, vox_tile_lists* out);
/* Where each input group starts in every tile's list: first holds nbounds
 * ascending triangle indices (the first triangle of each group, then the
 * total); out (host, ntiles * nbounds) receives, per tile, the position in
 * tile_tris of the tile's first triangle with index >= first[k].  Group g of
 * tile t is then tile_tris[out[t*nb + g] .. out[t*nb + g + 1]) (with
 * vox_add_group_gather, base = first[g]).  Not in the reference, whose tiler
 * keeps one list per texture (main.cu 3848-3880). */
int vox_tiler_group_bounds(vox_tiler* t, const uint32_t* first, uint32_t nbounds, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif
