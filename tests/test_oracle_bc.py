"""Pins the block-compression oracle (oracle/ref_bc.cpp) on the CPU: the BC1
and BC3 decoders on hand-decoded blocks of the published formats (RGB565
endpoints, 2-bit colour and 3-bit alpha indices, the 4- and 3-entry colour
modes, the 8- and 6-entry alpha modes), and the encoder through the decoder:
exact for uniform and two-valued blocks on representable values, every texel
on its nearest palette entry, edge texels repeated into padded blocks."""
import struct

import numpy as np

import oracle


def rgba(r, g, b, a=255):
    return r | (g << 8) | (b << 16) | (a << 24)


def ch(v, c):
    return (int(v) >> (8 * c)) & 255


def test_bc1_decode_four_colour_mode():
    idx = 0 | (1 << 2) | (2 << 4) | (3 << 6)
    blk = struct.pack("<HHI", 0xF800, 0x001F, idx)        # c0 red > c1 blue
    out = oracle.bc1_decode(blk)
    assert int(out[0]) == rgba(255, 0, 0)
    assert int(out[1]) == rgba(0, 0, 255)
    assert int(out[2]) == rgba(170, 0, 85)                 # (2a + b + 1) / 3
    assert int(out[3]) == rgba(85, 0, 170)
    assert all(int(v) == rgba(255, 0, 0) for v in out[4:])


def test_bc1_decode_three_colour_mode():
    idx = 0 | (1 << 2) | (2 << 4) | (3 << 6)
    blk = struct.pack("<HHI", 0x001F, 0xF800, idx)        # c0 <= c1
    out = oracle.bc1_decode(blk)
    assert int(out[2]) == rgba(127, 0, 127)                # (a + b) / 2
    assert int(out[3]) == 0                                # black, transparent
    # a BC3 colour block always has four entries
    out4 = oracle.bc1_decode(blk, four_always=True)
    assert int(out4[3]) & 0xFFFFFF == rgba(170, 0, 85) & 0xFFFFFF


def test_bc3_decode_alpha_modes():
    bits = 0
    for i in range(8):
        bits |= i << (3 * i)
    blk = bytes([255, 0]) + bits.to_bytes(6, "little") + struct.pack("<HHI", 0xFFFF, 0xFFFF, 0)
    out = oracle.bc3_decode(blk)
    assert [int(v) >> 24 for v in out[:8]] == [255, 0, 219, 182, 146, 109, 73, 36]
    blk = bytes([0, 255]) + bits.to_bytes(6, "little") + struct.pack("<HHI", 0, 0, 0)
    out = oracle.bc3_decode(blk)
    assert [int(v) >> 24 for v in out[:8]] == [0, 255, 51, 102, 153, 204, 0, 255]


def encode_block(tex16, nrm16=None):
    """One 4x4 brick slice block through the brick encoder (e = 4, 1 slice
    replicated 4 times in z)."""
    col = np.tile(np.asarray(tex16, np.uint32), 4)[None, :]
    nrm = np.tile(np.asarray(nrm16 if nrm16 is not None else tex16, np.uint32), 4)[None, :]
    c, n = oracle.bc_encode_bricks(col, nrm, (4, 4, 4))
    return c[:16].tobytes(), n[:8].tobytes()


def test_uniform_block_is_quantized_endpoint():
    v = rgba(201, 77, 13, 140)
    c, n = encode_block([v] * 16)
    out = oracle.bc3_decode(c)
    exp_r = ((201 * 31 + 127) // 255) << 3 | ((201 * 31 + 127) // 255) >> 2
    exp_g = ((77 * 63 + 127) // 255) << 2 | ((77 * 63 + 127) // 255) >> 4
    exp_b = ((13 * 31 + 127) // 255) << 3 | ((13 * 31 + 127) // 255) >> 2
    assert all(int(x) == rgba(exp_r, exp_g, exp_b, 140) for x in out)
    assert all(abs(ch(x, 0) - 201) <= 4 and abs(ch(x, 1) - 77) <= 2 for x in out)
    outn = oracle.bc1_decode(n)
    assert all(int(x) & 0xFFFFFF == rgba(exp_r, exp_g, exp_b) & 0xFFFFFF for x in outn)


def test_two_representable_colours_round_trip_exactly():
    a = rgba(255, 255, 255, 200)        # 565-exact values
    b = rgba(0, 0, 0, 10)
    rng = np.random.default_rng(1)
    tex = [a if x else b for x in rng.integers(0, 2, 16)]
    c, _ = encode_block(tex)
    out = oracle.bc3_decode(c)
    assert [int(x) for x in out] == [int(x) for x in tex]


def test_random_blocks_take_nearest_palette_entry():
    rng = np.random.default_rng(2)
    for _ in range(200):
        tex = [rgba(*map(int, rng.integers(0, 256, 4))) for _ in range(16)]
        c, _ = encode_block(tex)
        out = oracle.bc3_decode(c)
        a0, a1 = c[0], c[1]
        assert a0 == max(ch(t, 3) for t in tex) and a1 == min(ch(t, 3) for t in tex)
        # the four colour entries, decoded from index patterns
        c0, c1 = struct.unpack("<HH", c[8:12])
        assert c0 >= c1
        pal = [oracle.bc1_decode(struct.pack("<HHI", c0, c1, k * 0x55555555), True)[0]
               for k in range(4)]
        for t, d in zip(tex, out):
            dist = [sum((ch(t, q) - ch(p, q)) ** 2 for q in range(3)) for p in pal]
            got = sum((ch(t, q) - ch(d, q)) ** 2 for q in range(3))
            assert got == min(dist)
            alphas = sorted({a0, a1} | {((8 - k) * a0 + (k - 1) * a1 + 3) // 7 for k in range(2, 8)}) \
                if a0 != a1 else [a0]
            assert abs(ch(d, 3) - ch(t, 3)) == min(abs(ch(t, 3) - q) for q in alphas)


def test_brick_layout_and_edge_padding():
    e = 6                                           # B = 4: 2 x 2 blocks per slice
    rng = np.random.default_rng(3)
    col = rng.integers(0, 2 ** 32, size=(1, e ** 3), dtype=np.uint64).astype(np.uint32)
    nrm = (col & 0xFFFFFF).astype(np.uint32)
    c, n = oracle.bc_encode_bricks(col, nrm, (e, e, e))
    assert len(c) == 2 * 2 * e * 16 and len(n) == 2 * 2 * e * 8
    vol = col.reshape(e, e, e)
    for z in (0, 5):
        for by in range(2):
            for bx in range(2):
                k = (z * 2 + by) * 2 + bx
                out = oracle.bc3_decode(c[16 * k:16 * k + 16].tobytes()).reshape(4, 4)
                # padded texels decode like the edge texel they repeat
                for y in range(4):
                    for x in range(4):
                        sx, sy = min(bx * 4 + x, e - 1), min(by * 4 + y, e - 1)
                        ex_, ey_ = min(bx * 4 + x, e - 1) - bx * 4, min(by * 4 + y, e - 1) - by * 4
                        assert out[y, x] == out[ey_, ex_]
                # the block's alpha endpoints are the extremes of its texels
                blockv = [int(vol[z, min(by * 4 + y, e - 1), min(bx * 4 + x, e - 1)])
                          for y in range(4) for x in range(4)]
                assert c[16 * k] == max(v >> 24 for v in blockv)
                assert c[16 * k + 1] == min(v >> 24 for v in blockv)
