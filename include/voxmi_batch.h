/*
 * voxmi_batch.h -- many voxelizations in one hipGraph launch.
 *
 * Not in the reference, whose tiler calls the voxelizer once per tile with a
 * device synchronisation after every kernel (main.cu 4135-4405,
 * Voxelizer.cu 2216-2236).  Here a batch of voxelizer handles (meshes or
 * tiles, each with its own grid and brick size) is captured once into a
 * single hipGraph -- triangle setup, binning and the whole generate of every
 * job -- and replayed with one launch.  The voxelization proper of all jobs
 * is one persistent kernel working through every job's tiles from a device
 * queue (vox_batch_set_voxelize), so meshes of very different sizes share
 * the GPU in one launch rather than one small launch each.  The capture reuses the sizes an eager
 * vox_generate of each job measured, so there is no host round trip inside
 * the graph.  Before every launch the jobs' inputs are checksummed on the
 * device and their buffers compared with the captured ones: a launch on
 * changed inputs is refused (capture again), so a replay never runs with
 * sizes it was not built for.
 */
#ifndef VOXMI_BATCH_H
#define VOXMI_BATCH_H

#include <stdint.h>

#include "voxmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox_batch vox_batch;

typedef struct vox_batch_stats {
    int32_t  jobs;
    uint64_t launches;
    double   ms_graph_last;     /* HIP-event time of the last graph launch    */
    double   ms_check_last;     /* input checksums before it                  */
    double   ms_capture;        /* host time of the capture + instantiation   */
    uint64_t graph_nodes;
    uint32_t voxelize_launches; /* voxelize kernels in the graph               */
    uint32_t voxelize_blocks;   /* workgroups of each persistent voxelize launch
                                   (0: one launch per job and group)          */
} vox_batch_stats;

/* jobs: handles that have each completed one eager vox_generate with their
 * inputs uploaded (vox_allocate_triangle_memory).  The handles stay owned by
 * the caller and must outlive the batch. */
vox_batch* vox_batch_create(vox_ctx* const* jobs, int32_t njobs);
void vox_batch_destroy(vox_batch* b);
const char* vox_batch_get_error(vox_batch* b);
/* Capture every job's vox_compute_edges_face_normals_and_bounds +
 * vox_generate (no files) into one graph, then checksum the inputs. */
int vox_batch_capture(vox_batch* b);
/* Check the inputs and buffers against the capture, launch the graph, wait.
 * Each job's result is then available through vox_get_result. */
int vox_batch_launch(vox_batch* b);
int vox_batch_get_stats(vox_batch* b, vox_batch_stats* out);
/* How the next capture voxelizes.  The jobs chosen by `persistent` are
 * voxelized together: one persistent launch of `blocks` workgroups (0: 4 per
 * CU) per wave of groups (wave w holds the w-th group that overlaps its voxel
 * space of each of those jobs), its workgroups taking units of a few tiles
 * of any of them from a device queue, heaviest jobs first; the launch waits
 * for those jobs' tile maps and their colour passes wait for it.  The other
 * jobs voxelize their groups by launches of their own inside their branches.
 * persistent < 0 (the default): every job; 0: none; n > 0: the jobs whose
 * last generate found at most n active 8^3 tiles in level 0 (small jobs,
 * whose own launches would leave the GPU mostly idle).  Drops a captured
 * graph. */
int vox_batch_set_voxelize(vox_batch* b, int32_t persistent, uint32_t blocks);

#ifdef __cplusplus
}
#endif

#endif
