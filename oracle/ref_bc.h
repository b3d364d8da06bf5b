/*
 * ref_bc.h -- CPU restatement of the brick block compression (BC3 colours,
 * BC1 normals) and BC1 / BC3 decoders.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle for the HIP encoder; only tests/ load
 * it.
 */
#ifndef VOXMI_REF_BC_H
#define VOXMI_REF_BC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t ref_bc_blocks_per_brick(const uint32_t e[3]);
/* col / nrm: nbricks * e0*e1*e2 texels (x fastest); out_col 16 bytes and
 * out_nrm 8 bytes per block. */
void ref_bc_encode_bricks(const uint32_t* col, const uint32_t* nrm, uint64_t nbricks,
                          const uint32_t e[3], uint8_t* out_col, uint8_t* out_nrm);
/* four_always: interpret as a BC3 colour block (always 4 entries). */
void ref_bc1_decode(const uint8_t in[8], int four_always, uint32_t out[16]);
void ref_bc3_decode(const uint8_t in[16], uint32_t out[16]);

#ifdef __cplusplus
}
#endif

#endif
