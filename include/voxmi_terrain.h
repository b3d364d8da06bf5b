/*
 * voxmi_terrain.h -- C ABI of the procedural terrain on the device: the
 * seeded heightfield of the large paged workloads, built tile by tile where
 * the whole mesh would not fit (a 32769^2 grid is 2.1 G triangles).
 *
 * Not in the reference, which reads meshes from files through
 * OpenSceneGraph (main.cu 4518-4560); the generator is the scene generator
 * of this project's synthetic workloads (fBm of 2-D gradient noise, a
 * 256-texel height ramp), so a tile built here equals the host generator's
 * mesh on the same grid.  Every call runs on the null stream and returns when
 * its work is done, waiting for that work only (the library's other streams
 * are non-blocking): 1 ok, 0 error (vox_terrain_get_error, per thread).
 */
#ifndef VOXMI_TERRAIN_H
#define VOXMI_TERRAIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Heights of an n1 x n1 vertex grid over [0, extent]^2 into the device array
 * out (row j = y, x fastest): 0.5 E + 0.2 E fbm(x/E, y/E) with fBm of 6
 * octaves of gradient noise (lacunarity 2, gain 0.5, base frequency 4,
 * permutation = splitmix64(seed_fbm) Fisher-Yates), plus ridged_amp * E *
 * ridged(x/E, y/E), the normalised sum of (1 - |noise|)^2 over the same
 * octaves with its own permutation (seed_ridged); float64, rounded to
 * float32. */
int vox_terrain_heights(float* out, uint32_t n1, double extent, uint64_t seed_fbm,
                        uint64_t seed_ridged, double ridged_amp);
/* Height min and max over the vertex window [x0, x0 + w) x [y0, y0 + h). */
int vox_terrain_range(const float* heights, uint32_t n1, uint32_t x0, uint32_t y0, uint32_t w,
                      uint32_t h, float out_minmax[2]);
/* The triangles of cells [x0, x0 + cw) x [y0, y0 + ch), row by row, two per
 * cell, (v00, v10, v11) and (v00, v11, v01), into device arrays: verts (9
 * floats per triangle), vnormals (9: central differences in float32, the
 * grid's rim padded with its edge) and uvs (6: the texel centre of the
 * triangle's mean height on a tex_width x 1 ramp spanning [hmin, hmax], v =
 * 0.5). */
int vox_terrain_tile(const float* heights, uint32_t n1, double extent, uint32_t x0, uint32_t y0,
                     uint32_t cw, uint32_t ch, float hmin, float hmax, uint32_t tex_width,
                     float* verts, float* vnormals, float* uvs);
/* Heights at count host points (x, y pairs): the nearest grid vertex. */
int vox_terrain_sample(const float* heights, uint32_t n1, double extent, const float* xy,
                       uint32_t count, float* z);
const char* vox_terrain_get_error(void);

#ifdef __cplusplus
}
#endif

#endif
