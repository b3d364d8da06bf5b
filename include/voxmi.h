/*
 * voxmi.h -- C ABI of the MI355X voxelizer + GigaVoxels brick writer.
 *
 * Drop-in for the reference class cuda::Voxelizer
 * (apecoraro/VoxVizThesis VoxViz/Voxelizer/Voxelizer/Voxelizer.h 22-244),
 * which main.cu Voxelize() (3225-3341) drives as:
 *
 *   ctor -> initCuda -> allocateVoxelMemory -> setVoxelizationParams ->
 *   addTriangleGroup* -> allocateTriangleMemory ->
 *   computeEdgesFaceNormalsAndBounds -> generateVoxelsAndOctTree ->
 *   deallocateTriangleMemory  (resetCuda on failure)
 *
 * Each entry point below names the reference member it replaces.  Plain
 * pointers and sizes only; one handle per voxelizer; no global state.
 * Boolean members of the reference return 1 (true) / 0 (false) here;
 * vox_generate keeps the reference's 1 / 0 / -1 convention.  Error text
 * accumulates in the handle and is read with vox_get_error, as
 * Voxelizer::getErrorMessage (Voxelizer.cu 3176-3187) does.
 *
 * Semantics: the reference run with the chunk (voxChunkDim) the caller
 * passes.  A chunk of at least S = N + 4*((N/B)>>1) per axis holds the whole
 * voxel space in one pass.  A smaller chunk reproduces generateVoxelsAndOctTree's
 * chunk loop (Voxelizer.cu 2180-2814): chunk origins below N only (so the
 * last layers of S go unvoxelized when the chunks do not reach them), the
 * terrain fill stopping at each chunk's z = 0, a mip chain per chunk, node
 * samples clamped into each chunk with node types OR-ed over chunks and the
 * last chunk's sample as the constant colour, and bricks numbered in the
 * order the brick writer completes them (VoxelBrickWriter.cu 820-857,
 * 1026-1210).  Everything stays on the device either way.
 */
#ifndef VOXMI_H
#define VOXMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_ctx vox_ctx;

/* Texture address modes, numbered as cudaTextureAddressMode
 * (main.cu OpenGLModeToCudaMode 3207-3223). */
enum { VOX_ADDR_WRAP = 0, VOX_ADDR_CLAMP = 1, VOX_ADDR_MIRROR = 2, VOX_ADDR_BORDER = 3 };

/* Options (vox_set_options). */
enum {
    /* Build mip level d from level d-1.  Default (flag clear) reproduces the
     * reference's computeMipMaps recursion, where level d >= 2 reads level
     * d-2 (Voxelizer.cu 513-517). */
    VOX_OPT_MIP_CHAIN = 1u << 0,
    /* Reserved: every mip level's sparse tiles are kept until the next
     * vox_generate, so vox_copy_level always works. */
    VOX_OPT_KEEP_LEVELS = 1u << 1,
    /* Skip the dense S^3 occupancy bitmap (it is produced by default). */
    VOX_OPT_NO_OCCUPANCY = 1u << 2,
    /* Record per-kernel HIP-event timings for vox_get_stats. */
    VOX_OPT_TIMING = 1u << 3
};

/* One triangle group: the arguments of cuda::Voxelizer::addTriangleGroup
 * (Voxelizer.h 128-162).  Triangle soup, 3 vertices per triangle.
 * vnormals, uvs, rgba8 may be NULL.  A group is textured when both uvs and
 * rgba8 are given (ComputeColorsAndNormals); otherwise it is coloured by
 * its normals (ComputeColorsAsNormals). */
typedef struct vox_group {
    const float*   verts;      /* 9*ntri floats                               */
    const float*   vnormals;   /* 9*ntri floats or NULL                        */
    const float*   uvs;        /* 6*ntri floats or NULL                        */
    uint64_t       ntri;
    float          bbox_min[3];/* group bounds, world units                    */
    float          bbox_max[3];
    int32_t        is_terrain; /* fill up to 19 voxels below textured hits     */
    const uint8_t* rgba8;      /* width*height*4 bytes or NULL                 */
    int32_t        width, height;
    int32_t        addr0, addr1;
} vox_group;

/* Device-resident result of the last vox_generate.  All pointers are HIP
 * device pointers owned by the handle; valid until the next vox_generate,
 * vox_reset or vox_destroy. */
#define VOX_MAX_LEVELS 16
typedef struct vox_result {
    int32_t  levels;                 /* octree depth count L                   */
    uint32_t S[3];                   /* voxel space per axis                   */
    uint32_t brick_dim[3];           /* B                                      */
    const uint32_t* occupancy;       /* S^3 bits, bit i = voxel (z*Sy+y)*Sx+x  */
    uint64_t occupancy_words;
    /* per tree depth d (0 = root): nodes = 8^d, x fastest, then y, z */
    uint64_t        nodes[VOX_MAX_LEVELS];
    const uint8_t*  types[VOX_MAX_LEVELS];       /* 0 CONST,1 NON-CONST,2 LEAF-CONST */
    const uint32_t* const_colors[VOX_MAX_LEVELS];/* RGBA8, R in the low byte    */
    const uint32_t* brick_index[VOX_MAX_LEVELS]; /* 0xffffffff when no brick    */
    uint64_t        bricks[VOX_MAX_LEVELS];      /* NON-CONST nodes at depth d  */
    const uint32_t* brick_colors[VOX_MAX_LEVELS];/* bricks*(B+2)^3 RGBA8        */
    const uint32_t* brick_normals[VOX_MAX_LEVELS];/* bricks*(B+2)^3 RGB8 (+0)   */
    const uint32_t* brick_node[VOX_MAX_LEVELS];  /* node index of each brick    */
} vox_result;

/* Work counts and timings of the last vox_generate. */
typedef struct vox_stats {
    uint64_t triangles;               /* T                                      */
    uint64_t pairs;                   /* (triangle, tile) bin entries           */
    uint64_t sat_tests;               /* (triangle, voxel) overlap tests        */
    uint64_t active_tiles[VOX_MAX_LEVELS]; /* 8^3 voxel tiles per mip level    */
    uint64_t occupied_voxels;
    uint64_t bricks_total;
    double   ms_total;                /* HIP-event time of the device pipeline  */
    int32_t  nkernels;
    char     kernel_name[64][32];
    double   kernel_ms[64];           /* summed per kernel name                */
    uint32_t kernel_launches[64];
    /* With VOX_OPT_TIMING, one voxel count per mip level k >= 1: the stored
     * source voxels its filter reads (active source tiles under its active
     * tiles; on level 0 only the voxels some group hit or a fill reached).
     * The algorithmic bytes of the filter are 16 per such voxel plus 20 per
     * voxel of its active tiles.  Zero with several chunks. */
    uint64_t mip_source_voxels[VOX_MAX_LEVELS];
    uint64_t bins;                    /* (group, tile) bins of the setup pass  */
    uint64_t scan_values;             /* values the generate's prefix sums read */
} vox_stats;

/* cuda::Voxelizer::Voxelizer(voxDim, brickDim, voxChunkDim)
 * (Voxelizer.h 122-124, Voxelizer.cu 125-153).  Always returns a handle
 * (NULL only when out of host memory); invalid dimensions (N/B not a power
 * of two >= 2, unequal ratios, a chunk below S that is not a multiple of
 * 2^(levels-1), more than 64 chunks per axis) are reported by vox_get_error
 * and make every later call fail.  chunk_dim may be NULL: one chunk covering
 * the voxel space.  A NON-CONST node whose brick was never completed (a
 * chunk smaller than S that leaves part of its border unvoxelized) has brick
 * index 0xffffffff in vox_result (written as 0 in tree.gvx, as the reference
 * does). */
vox_ctx* vox_create(const uint32_t vox_dim[3], const uint32_t brick_dim[3],
                    const uint32_t chunk_dim[3]);
/* cuda::Voxelizer::~Voxelizer (Voxelizer.cu 272-277). */
void vox_destroy(vox_ctx* h);

/* cuda::Voxelizer::setVoxelizationParams (Voxelizer.h 164, Voxelizer.cu 155-161). */
int vox_set_params(vox_ctx* h, const float p[3], const float delta[3]);
/* Options bit set (VOX_OPT_*); not in the reference. */
int vox_set_options(vox_ctx* h, uint32_t options);
/* Launch shape: workgroups of the persistent voxelize and colour kernels
 * (default 8 per compute unit); the autotuner's knob.  Not in the
 * reference. */
int vox_set_launch(vox_ctx* h, uint32_t persistent_blocks);

/* cuda::Voxelizer::addTriangleGroup (Voxelizer.h 128-162).  Host pointers;
 * like the reference, they are only read by vox_allocate_triangle_memory. */
int vox_add_group(vox_ctx* h, const vox_group* g);
/* Same, but every pointer in *g is a device pointer already resident in HBM
 * (no copy is made; the caller keeps them alive until vox_generate returns). */
int vox_add_group_device(vox_ctx* h, const vox_group* g);
/* A group of n triangles picked from device-resident arrays: triangle k is
 * row idx[k] - base of g's verts / vnormals / uvs (device pointers of the
 * whole group, g->ntri rows; idx is a device array whose values lie in
 * [base, base + g->ntri)).  The rows are gathered on the device by
 * vox_allocate_triangle_memory, which also sets the group's bounds to those
 * of the picked triangles, as the reference's tiler does for the triangles
 * of a tile (main.cu 3848-3880); g->bbox is ignored.  For tiles voxelized
 * from a scene resident in HBM (with the tiler's per-tile lists).  Not in
 * the reference. */
int vox_add_group_gather(vox_ctx* h, const vox_group* g, const uint32_t* idx, uint64_t n,
                         uint32_t base);

/* cuda::Voxelizer::initCuda (Voxelizer.cu 297-322): device 0, own stream. */
int vox_init(vox_ctx* h);
/* cuda::Voxelizer::resetCuda (Voxelizer.cu 324-344): frees everything,
 * recreates the handle's stream and events and reallocates voxel memory.
 * Unlike the reference it does not reset the device, which other users of
 * the process may share. */
int vox_reset(vox_ctx* h);
/* cuda::Voxelizer::allocateVoxelMemory / deallocateVoxelMemory
 * (Voxelizer.cu 163-236). */
int  vox_allocate_voxel_memory(vox_ctx* h);
void vox_deallocate_voxel_memory(vox_ctx* h);
/* cuda::Voxelizer::allocateTriangleMemory / deallocateTriangleMemory
 * (Voxelizer.cu 346-457, 238-242): H2D of host groups. */
int  vox_allocate_triangle_memory(vox_ctx* h);
void vox_deallocate_triangle_memory(vox_ctx* h);
/* cuda::Voxelizer::computeEdgesFaceNormalsAndBounds (Voxelizer.cu 459-489):
 * per-triangle setup and binning of triangles to 8^3 voxel tiles. */
int vox_compute_edges_face_normals_and_bounds(vox_ctx* h);
/* cuda::Voxelizer::generateVoxelsAndOctTree (Voxelizer.cu 2180-2814).
 * Returns 1 ok, 0 error, -1 no group overlaps the grid.  out_dir may be NULL
 * (results stay in HBM, see vox_get_result); otherwise tree.gvx and
 * voxels_<depth>.gvb are written there.  compressed = 1 also produces
 * block-compressed bricks (BC3 colours, BC1 normals, see vox_get_compressed)
 * and writes them to the .gvb files, as WriteCompressedBinaryBrick does
 * (VoxelBrickWriter.cu 590-723); the reference had the OpenGL driver encode
 * them, here the project's deterministic encoder does. */
int vox_generate(vox_ctx* h, const char* out_dir, int binary, int compressed);
/* cuda::Voxelizer::getErrorMessage (Voxelizer.cu 3176-3187). */
const char* vox_get_error(vox_ctx* h);

/* In-memory result of the last vox_generate (device pointers). */
int vox_get_result(vox_ctx* h, vox_result* out);
/* Work counts and timings of the last vox_generate. */
int vox_get_stats(vox_ctx* h, vox_stats* out);
/* Copy n bytes from a device pointer of the result to host memory. */
int vox_copy_to_host(vox_ctx* h, void* dst, const void* src_device, uint64_t nbytes);
/* Materialise mip level k (0 = full resolution) densely on the host:
 * colors (RGBA8) and normals (3 floats per voxel), x fastest; dims receives
 * the level's extent.  With one chunk the extent is S_k = S >> k.  With
 * several chunks each chunk has its own mip chain and the level is laid out
 * chunk by chunk: (chunks per axis) * (chunk >> k) voxels per axis, chunk i
 * of an axis at offset i * (chunk >> k) (level 0: the voxel space itself,
 * cut or zero-padded to that extent).  colors and normals both NULL: only
 * dims is written. */
int vox_copy_level(vox_ctx* h, int32_t level, uint32_t* colors, float* normals,
                   uint32_t dims[3]);
/* Block-compressed bricks of the last vox_generate(..., compressed = 1):
 * per depth, bricks * blocks_per_brick blocks, blocks_per_brick =
 * ceil((B+2)/4)^2 * (B+2) (each slice of a brick in 4x4 blocks, row-major,
 * slices in z).  Colours: 16-byte BC3 blocks; normals: 8-byte BC1 blocks.
 * Returns 0 when the last generate was not compressed.  Not in the
 * reference (which kept compressed bricks only in files). */
typedef struct vox_compressed {
    int32_t        levels;
    uint64_t       blocks_per_brick;
    uint64_t       bricks[VOX_MAX_LEVELS];
    const uint8_t* colors[VOX_MAX_LEVELS];
    const uint8_t* normals[VOX_MAX_LEVELS];
} vox_compressed;
int vox_get_compressed(vox_ctx* h, vox_compressed* out);

/* Asynchronous generate, not in the reference (whose pager thread loads
 * finished tiles from disk while frames go on, GigaVoxelsDatabasePager.cpp
 * 232-338): vox_generate(h, NULL, 1, compressed) -- with the setup pass if
 * it has not run -- on a host worker thread, so the caller can go on issuing
 * work for other handles (ray casts, other tiles) while this one's device
 * work and host round trips proceed on its own stream.  Until vox_wait has
 * returned its result every other call on h fails (vox_get_error says so),
 * except vox_destroy, which waits for the worker first.  Returns 1 when the
 * worker started. */
int vox_generate_async(vox_ctx* h, int32_t compressed);
/* Result of the running or finished vox_generate_async: 1 ok, 0 error,
 * -1 no group overlaps the grid, 2 still running (only when block = 0). */
int vox_wait(vox_ctx* h, int32_t block);
/* Free what only the next vox_generate needs (triangles, bins, level tiles
 * and their work lists), keeping the result (vox_get_result and
 * vox_get_compressed stay valid; vox_copy_level does not).  For callers
 * that keep many results resident, such as a pager of tiles.  The next
 * vox_generate needs its groups added and uploaded again.  Not in the
 * reference. */
int vox_release_work_memory(vox_ctx* h);

/* Device memory for callers with no GPU array library of their own (a scene
 * kept resident in HBM for vox_add_group_gather or vox_add_group_device):
 * vox_device_alloc returns NULL on failure; vox_device_upload copies host
 * bytes to dst + offset and returns when they are on the device; free
 * waits for the device first.  Not in the reference. */
void* vox_device_alloc(uint64_t bytes);
int vox_device_upload(void* dst, uint64_t offset, const void* src, uint64_t bytes);
int vox_device_free(void* p);
/* Device-to-device copy by a plain copy kernel (16-byte aligned pointers, a
 * multiple of 16 bytes); returns when done.  Its known traffic is the
 * reference point that calibrates hardware traffic counters. */
int vox_device_copy(void* dst, const void* src, uint64_t bytes);
/* The same copy with its 256-byte blocks visited in a scattered order (block
 * i of the launch copies block (i * 0x9E3779B1) mod n, n = bytes / 256 a
 * power of two): every byte still read and written once, but neighbouring
 * workgroups touch distant lines, as a gathering kernel does.  A second
 * reference point for the traffic counters.  Not in the reference. */
int vox_device_copy_scattered(void* dst, const void* src, uint64_t bytes);

/* Copy bytes from src + offset (device) to host memory dst; returns when
 * they are on the host.  The counterpart of vox_device_upload.  Not in the
 * reference. */
int vox_device_download(void* dst, const void* src, uint64_t offset, uint64_t bytes);

/* Exclusive prefix sum of n uint32 values on the device, modulo 2^32:
 * out[i] = in[0] + ... + in[i-1] for i <= n (out holds n + 1 values, out[n]
 * the total).  in and out are device pointers, 16-byte aligned, not
 * overlapping; clear_input = 1 also sets in[0, n) to zero (counters that
 * count again).  It is the single-pass decoupled look-back scan that sizes
 * the voxelizer's (group, tile) bins, tile lists and brick addresses; its
 * look-back scratch is allocated and freed by the call.  Runs on the null
 * stream and returns when done: 1 ok, 0 error (vox_scan_get_error, per
 * calling thread).  Not in the reference, whose triangle batches are fixed
 * 512-triangle launches (Voxelizer.cu 2359-2400) with no prefix sum. */
int vox_scan_u32(uint32_t* in, uint32_t* out, uint64_t n, int32_t clear_input);

/* vox_scan_u32 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as vox_scan_u32, for n up to 2^32
 * values (larger n is rejected with an error before any device work).  The
 * look-back scratch is a fixed 4 MiB device array inside the library, one
 * per GPU, held until the process ends; each call clears the part it uses.
 * Calls on one GPU share that array, so calls that overlap in time on one
 * GPU (from several host threads) race: serialise them. */
int vox_scan_u32_noalloc(uint32_t* in, uint32_t* out, uint64_t n, int32_t clear_input);
const char* vox_scan_get_error(void);

/* Write tree.gvx and voxels_<depth>.gvb of the last vox_generate into
 * out_dir (host-side writer, Voxelizer.cu 2816-3061, VoxelBrickWriter.cu
 * 725-818). */
int vox_write_files(vox_ctx* h, const char* out_dir);

#ifdef __cplusplus
}
#endif

#endif
