/*
 * ref_raycast.h -- CPU restatement of the GigaVoxels request-producing ray
 * cast (VoxViz/Shaders/GigaVoxels.frag) with the renderer's uniforms
 * (VoxViz/GigaVoxels/GigaVoxelsRenderer.cpp 1650-1816).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle for the HIP ray cast.  Only tests/
 * load it; the product never links or calls it.
 */
#ifndef VOXMI_REF_RAYCAST_H
#define VOXMI_REF_RAYCAST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_ray_camera {
    float    position[3], look[3], up[3];
    float    fov_y_deg;
    int32_t  width, height;
    uint32_t frame;
} ref_ray_camera;

typedef struct ref_ray_params {
    float    vol_min[3], vol_max[3];
    float    ray_step, lod_scalar;
    uint32_t jitter_seed, max_traversals;
} ref_ray_params;

/* Shades every pixel (row 0 at the bottom): usage_out receives 12 node ids
 * per pixel, rgba_out (may be NULL) the composited colour.  nodes / pool are
 * the node texture (2 words per texel) and the brick pool colour texture.
 * Returns 0 for an empty volume box. */
int32_t ref_raycast(const ref_ray_camera* cam, const ref_ray_params* prm, int32_t levels,
                    const uint32_t B[3], const uint32_t* nodes, const uint32_t ndim[3],
                    const uint32_t* pool, const uint32_t pdim[3], uint32_t* usage_out,
                    uint32_t* rgba_out);

/* The same over a wide brick pool (vox_brick_pool_create_wide) when wide is
 * non-zero: pdim = (B+2, B+2, (B+2) * slots), word 1 of a resident node is
 * the slot, and each brick is sampled in its own (B+2)^3 texel frame. */
int32_t ref_raycast_pool(const ref_ray_camera* cam, const ref_ray_params* prm, int32_t levels,
                         const uint32_t B[3], const uint32_t* nodes, const uint32_t ndim[3],
                         const uint32_t* pool, const uint32_t pdim[3], int32_t wide,
                         uint32_t* usage_out, uint32_t* rgba_out);

#ifdef __cplusplus
}
#endif

#endif
