"""The premise of the chunked mip's single-chunk shortcut (CPU).

With several chunks, level voxel X of chunk i = X // Cd reads source voxels
i * Cs + 2x and i * Cs + min(2x + 1, Cs - 1), x = X - i * Cd (the per-chunk
mip recursion, VoxelizerKernels.cu 1520-1663 run per chunk by
Voxelizer.cu 491-594).  Where every axis's Cd is a multiple of the 8-voxel
tile, Cs - 2 Cd a multiple of 16 and the last voxel's second source below
the source level's end, the library filters the level as one chunk with the
source offset i * (Cs - 2 Cd) added per tile.  This checks, for the chunk
configurations the GPU tests and the bench use, that the rule picks exactly
those levels and that under it every voxel's two sources are 2X + o and
2X + o + 1, inside the source level, in the tile pair 2t + o / 8 that the
shortcut reads, with the parities it assumes.
"""
import pytest

TILE = 8


def geometry(N, B, chunk):
    r = N // B
    L = 1
    while (B << (L - 1)) != N:
        L += 1
    S = N + 4 * (r >> 1)
    if chunk >= S:
        C, nc = S, 1
    else:
        C, nc = chunk, (N + chunk - 1) // chunk
    dims = [S if k == 0 else nc * (C >> k) for k in range(L)]
    return L, C, dims


def shift_rule(C, dims, ks, kd):
    Cd, Cs = C >> kd, C >> ks
    if Cd == 0 or Cd % TILE or Cs < 2 * Cd or (Cs - 2 * Cd) % (2 * TILE) or dims[kd] == 0:
        return None
    X = dims[kd] - 1
    i, x = X // Cd, X % Cd
    if i * Cs + 2 * x + 1 >= dims[ks]:
        return None
    return Cd, Cs - 2 * Cd


CASES = [(32, 4, 16, False), (64, 8, 16, False), (64, 8, 32, False), (64, 8, 48, False),
         (128, 8, 64, False), (256, 8, 64, True), (256, 8, 128, False), (256, 8, 256, False),
         (512, 16, 256, False), (512, 16, 320, False), (1024, 16, 256, False),
         (1024, 16, 256, True)]


@pytest.mark.parametrize("N,B,chunk,chain", CASES)
def test_shift_rule_matches_the_per_chunk_sources(N, B, chunk, chain):
    L, C, dims = geometry(N, B, chunk)
    for kd in range(1, L):
        ks = kd - 1 if chain else (0 if kd == 1 else kd - 2)
        rule = shift_rule(C, dims, ks, kd)
        if rule is None:
            continue
        Cd, off = rule
        Cs = C >> ks
        for X in range(dims[kd]):
            i, x = X // Cd, X % Cd
            s0, s1 = i * Cs + 2 * x, i * Cs + min(2 * x + 1, Cs - 1)
            o = (X // TILE * TILE) // Cd * off          # the tile's offset
            assert (s0, s1) == (2 * X + o, 2 * X + o + 1), (kd, X)
            assert s1 < dims[ks]
            t = X // TILE
            assert s0 // TILE in (2 * t + o // TILE, 2 * t + o // TILE + 1)
            assert (s0 // TILE) % 2 == ((2 * X) // TILE) % 2


def test_rule_takes_the_fine_levels_of_the_reference_default():
    # the 1024^3 / 16^3 workload at the reference's chunk of 256: levels 1-5
    # take the shortcut, level 6 (4 voxels per chunk) does not
    L, C, dims = geometry(1024, 16, 256)
    took = [kd for kd in range(1, L)
            if shift_rule(C, dims, 0 if kd == 1 else kd - 2, kd) is not None]
    assert took == [1, 2, 3, 4, 5]
    # 320: level 1's sources of the second chunk would pass the level-0 end
    L, C, dims = geometry(512, 16, 320)
    assert shift_rule(C, dims, 0, 1) is None
