/*
 * ref_cpu.h -- CPU restatement of the reference voxelizer + GigaVoxels brick
 * writer (apecoraro/VoxVizThesis, VoxViz/Voxelizer/Voxelizer).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * The restatement follows the reference's own data flow: dense per-voxel
 * arrays for the whole chunk, a per-triangle bounding-box loop for the
 * overlap test (the shape of the commented-out cpuTest, VoxelizerKernels.cu
 * 2382-2509), a dense colour pass per triangle group, the recursive mip
 * chain of Voxelizer.cu 491-594, the per-voxel node typing kernels of
 * VoxelizerKernels.cu 1680-2380 (including the two-launch x split of
 * Voxelizer.cu 1081-1120), the brick gather of VoxelBrickWriter.cu 725-818
 * and the branch collapse of Voxelizer.cu 701-774.
 *
 * fp32 throughout, compiled with -ffp-contract=off, GLM 0.9.3 operation
 * order (dot = (x*x + y*y) + z*z, normalize = v * (1/sqrt(dot))).
 */
#ifndef VOXMI_REF_CPU_H
#define VOXMI_REF_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One triangle group, as handed to cuda::Voxelizer::addTriangleGroup
 * (Voxelizer.h 128-162).  Triangle soup: 3 vertices per triangle. */
typedef struct ref_group {
    const float*   verts;      /* 9*ntri floats (x,y,z per vertex)            */
    const float*   vnormals;   /* 9*ntri floats or NULL (face normals used)   */
    const float*   uvs;        /* 6*ntri floats or NULL                        */
    uint64_t       ntri;
    float          bbox_min[3];/* group bounds in world units                 */
    float          bbox_max[3];
    int32_t        is_terrain;
    const uint8_t* rgba8;      /* w*h*4 bytes or NULL (untextured group)      */
    int32_t        width, height;
    int32_t        addr0, addr1; /* 0 wrap, 1 clamp, 2 mirror, 3 border      */
} ref_group;

typedef struct ref_result ref_result;

/* flags bit 0 (REF_MIP_CHAIN): level d reads level d-1.  Clear: the
 * reference's live recursion, where level d>=2 reads level d-2
 * (Voxelizer.cu 513-517).
 * flags bit 1 (REF_FP64_TWIN): also run the fp64 twin of the colour/normal
 * pass (see ref_normal_error_fp64).
 * status receives 1 (ok), 0 (error), -1 (no group overlaps the grid),
 * mirroring generateVoxelsAndOctTree (Voxelizer.cu 2798-2813). */
#define REF_MIP_CHAIN 1
#define REF_FP64_TWIN 2
ref_result* ref_run(const uint32_t N[3], const uint32_t B[3],
                    const float p[3], const float delta[3],
                    const ref_group* groups, int32_t ngroups,
                    int32_t flags, int32_t* status);
const char* ref_error(void);

/* The same run with the voxel space processed in chunks of chunk[a] voxels
 * per axis, as generateVoxelsAndOctTree does (Voxelizer.cu 2180-2814): chunks
 * visited breadth first from the origin, each voxelized, coloured, filled
 * and mip-mapped on its own (the terrain fill stops at the chunk's z = 0),
 * node samples clamped into the chunk and node types OR-ed over chunks (the
 * constant colour is the last chunk's sample), bricks that reach a chunk's
 * edge assembled piecewise and written once complete (VoxelBrickWriter.cu
 * 126-205, 1026-1210), and brick indices in that write order.  A chunk at
 * least N + 4*((N/B)>>1) wide is the single-chunk run (chunk NULL: the same);
 * a smaller one must be a multiple of 2^(levels-1).  Levels are returned
 * chunk by chunk: level k is (chunks per axis) * (chunk >> k) voxels per
 * axis, chunk (i, j, l) at offset (i, j, l) * (chunk >> k).  A NON-CONST node
 * whose brick was never completed keeps brick index 0xffffffff (the
 * reference's tree file writes 0 for it, VoxelBrickWriter.h 143-149). */
ref_result* ref_run_chunked(const uint32_t N[3], const uint32_t B[3], const uint32_t chunk[3],
                            const float p[3], const float delta[3],
                            const ref_group* groups, int32_t ngroups,
                            int32_t flags, int32_t* status);

int32_t         ref_levels(const ref_result* r);
void            ref_dims(const ref_result* r, uint32_t S[3]);
const uint32_t* ref_occupancy(const ref_result* r, uint64_t* nwords);
const uint8_t*  ref_types(const ref_result* r, int32_t depth, uint64_t* n);
const uint32_t* ref_const_colors(const ref_result* r, int32_t depth, uint64_t* n);
const uint32_t* ref_brick_index(const ref_result* r, int32_t depth, uint64_t* n);
uint64_t        ref_num_bricks(const ref_result* r, int32_t depth);
const uint32_t* ref_brick_colors(const ref_result* r, int32_t depth);
const uint32_t* ref_brick_normals(const ref_result* r, int32_t depth);
const uint32_t* ref_level_colors(const ref_result* r, int32_t level, uint32_t dims[3]);
const float*    ref_level_normals(const ref_result* r, int32_t level, uint32_t dims[3]);
void            ref_free(ref_result* r);

/* fp64 twin of the colour/normal pass for the leaf level: same hit sets and
 * slots as the fp32 run, normals accumulated and normalised in double.
 * Returns the largest angle (radians) between the fp32 and fp64 unit normals
 * over voxels whose fp32 normal is finite and non-zero, and the count. */
double ref_normal_error_fp64(const ref_result* r, uint64_t* nvox);

#ifdef __cplusplus
}
#endif

#endif
