// ref_bc.cpp -- CPU restatement of the brick block compression used for
// compressed .gvb output (BC3 colours, BC1 normals), with decoders.
//
// TEST INFRASTRUCTURE ONLY: the oracle for the HIP encoder; only tests/ load
// it.  The reference had the OpenGL driver encode (VoxelBrickWriter.cu
// 590-723); the encoder restated here is the project's own: per-channel
// max / min endpoints quantized to RGB565 with rounding, c0 >= c1, palette
// of the expanded endpoints and their rounded thirds, nearest entry in
// squared RGB distance with the lowest index on ties; alpha endpoints max /
// min with the 8-entry interpolation ((8-k)a0 + (k-1)a1 + 3) / 7.  Bricks of
// e^3 texels are compressed slice by slice in 4x4 blocks, row-major, slices in
// z, texels past the edge repeating the edge.
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "ref_bc.h"

namespace {

int quant5(int v) { return (v * 31 + 127) / 255; }
int quant6(int v) { return (v * 63 + 127) / 255; }
int expand5(int q) { return (q << 3) | (q >> 2); }
int expand6(int q) { return (q << 2) | (q >> 4); }

void palette4(uint16_t c0, uint16_t c1, bool four, int P[4][3]) {
    P[0][0] = expand5(c0 >> 11); P[0][1] = expand6((c0 >> 5) & 63); P[0][2] = expand5(c0 & 31);
    P[1][0] = expand5(c1 >> 11); P[1][1] = expand6((c1 >> 5) & 63); P[1][2] = expand5(c1 & 31);
    for (int c = 0; c < 3; ++c) {
        if (four) {
            P[2][c] = (2 * P[0][c] + P[1][c] + 1) / 3;
            P[3][c] = (P[0][c] + 2 * P[1][c] + 1) / 3;
        } else {
            P[2][c] = (P[0][c] + P[1][c]) / 2;
            P[3][c] = 0;
        }
    }
}

// texels: 16 RGB(A) words, byte 0 = R
void color_block(const uint32_t* t, uint8_t out[8]) {
    int mn[3] = {255, 255, 255}, mx[3] = {0, 0, 0};
    for (int i = 0; i < 16; ++i)
        for (int c = 0; c < 3; ++c) {
            int v = (t[i] >> (8 * c)) & 255;
            if (v < mn[c]) mn[c] = v;
            if (v > mx[c]) mx[c] = v;
        }
    uint16_t c0 = (uint16_t)((quant5(mx[0]) << 11) | (quant6(mx[1]) << 5) | quant5(mx[2]));
    uint16_t c1 = (uint16_t)((quant5(mn[0]) << 11) | (quant6(mn[1]) << 5) | quant5(mn[2]));
    if (c0 < c1) {
        uint16_t s = c0;
        c0 = c1;
        c1 = s;
    }
    uint32_t idx = 0;
    if (c0 != c1) {
        int P[4][3];
        palette4(c0, c1, true, P);
        for (int i = 0; i < 16; ++i) {
            int best = 1 << 30, bi = 0;
            for (int k = 0; k < 4; ++k) {
                int d = 0;
                for (int c = 0; c < 3; ++c) {
                    const int e = (int)((t[i] >> (8 * c)) & 255) - P[k][c];
                    d += e * e;
                }
                if (d < best) {
                    best = d;
                    bi = k;
                }
            }
            idx |= (uint32_t)bi << (2 * i);
        }
    }
    std::memcpy(out, &c0, 2);
    std::memcpy(out + 2, &c1, 2);
    std::memcpy(out + 4, &idx, 4);
}

void alpha_block(const uint32_t* t, uint8_t out[8]) {
    int a0 = 0, a1 = 255;
    for (int i = 0; i < 16; ++i) {
        const int a = (int)(t[i] >> 24);
        if (a > a0) a0 = a;
        if (a < a1) a1 = a;
    }
    uint64_t bits = 0;
    if (a0 != a1) {
        int A[8] = {a0, a1, 0, 0, 0, 0, 0, 0};
        for (int k = 2; k < 8; ++k) A[k] = ((8 - k) * a0 + (k - 1) * a1 + 3) / 7;
        for (int i = 0; i < 16; ++i) {
            const int a = (int)(t[i] >> 24);
            int best = 1 << 30, bi = 0;
            for (int k = 0; k < 8; ++k) {
                const int d = std::abs(a - A[k]);
                if (d < best) {
                    best = d;
                    bi = k;
                }
            }
            bits |= (uint64_t)bi << (3 * i);
        }
    }
    out[0] = (uint8_t)a0;
    out[1] = (uint8_t)a1;
    for (int b = 0; b < 6; ++b) out[2 + b] = (uint8_t)(bits >> (8 * b));
}

void gather(const uint32_t* slice, uint32_t ex, uint32_t ey, uint32_t bx, uint32_t by,
            uint32_t t[16]) {
    for (int i = 0; i < 16; ++i) {
        uint32_t x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
        if (x >= ex) x = ex - 1;
        if (y >= ey) y = ey - 1;
        t[i] = slice[(uint64_t)y * ex + x];
    }
}

}  // namespace

extern "C" uint64_t ref_bc_blocks_per_brick(const uint32_t e[3]) {
    return (uint64_t)((e[0] + 3) / 4) * ((e[1] + 3) / 4) * e[2];
}

extern "C" void ref_bc_encode_bricks(const uint32_t* col, const uint32_t* nrm, uint64_t nbricks,
                                     const uint32_t e[3], uint8_t* out_col, uint8_t* out_nrm) {
    const uint32_t bxn = (e[0] + 3) / 4, byn = (e[1] + 3) / 4;
    const uint64_t nv = (uint64_t)e[0] * e[1] * e[2];
    uint64_t k = 0;
    for (uint64_t b = 0; b < nbricks; ++b)
        for (uint32_t z = 0; z < e[2]; ++z)
            for (uint32_t by = 0; by < byn; ++by)
                for (uint32_t bx = 0; bx < bxn; ++bx, ++k) {
                    uint32_t t[16];
                    gather(col + b * nv + (uint64_t)z * e[0] * e[1], e[0], e[1], bx, by, t);
                    alpha_block(t, out_col + 16 * k);
                    color_block(t, out_col + 16 * k + 8);
                    gather(nrm + b * nv + (uint64_t)z * e[0] * e[1], e[0], e[1], bx, by, t);
                    color_block(t, out_nrm + 8 * k);
                }
}

extern "C" void ref_bc1_decode(const uint8_t in[8], int four_always, uint32_t out[16]) {
    uint16_t c0, c1;
    uint32_t idx;
    std::memcpy(&c0, in, 2);
    std::memcpy(&c1, in + 2, 2);
    std::memcpy(&idx, in + 4, 4);
    int P[4][3];
    palette4(c0, c1, four_always || c0 > c1, P);
    for (int i = 0; i < 16; ++i) {
        const int k = (idx >> (2 * i)) & 3;
        out[i] = (uint32_t)P[k][0] | ((uint32_t)P[k][1] << 8) | ((uint32_t)P[k][2] << 16) |
                 ((!four_always && c0 <= c1 && k == 3) ? 0u : 0xff000000u);
    }
}

extern "C" void ref_bc3_decode(const uint8_t in[16], uint32_t out[16]) {
    const int a0 = in[0], a1 = in[1];
    uint64_t bits = 0;
    for (int b = 0; b < 6; ++b) bits |= (uint64_t)in[2 + b] << (8 * b);
    int A[8] = {a0, a1, 0, 0, 0, 0, 0, 0};
    if (a0 > a1) {
        for (int k = 2; k < 8; ++k) A[k] = ((8 - k) * a0 + (k - 1) * a1 + 3) / 7;
    } else {
        for (int k = 2; k < 6; ++k) A[k] = ((6 - k) * a0 + (k - 1) * a1 + 2) / 5;
        A[6] = 0;
        A[7] = 255;
    }
    ref_bc1_decode(in + 8, 1, out);
    for (int i = 0; i < 16; ++i)
        out[i] = (out[i] & 0x00ffffffu) | ((uint32_t)A[(bits >> (3 * i)) & 7] << 24);
}
