"""HIP path against the oracle with chunks smaller than the voxel space (GPU).

generateVoxelsAndOctTree processes the voxel space in chunks of voxChunkDim
(Voxelizer.cu 2180-2814; main.cu keeps the default of 256).  Every case here
compares occupancy, every level (laid out chunk by chunk), node types,
constant colours, brick indices and brick payloads bit-exactly with the
oracle's chunked run: seams inside bricks at every level, bricks assembled
across three chunks per axis, layers the chunks never reach (chunk origins
stay below N), chunks wider than the voxel space's remainder, the terrain
fill cut at chunk seams, and the headline workload's own generator (terrain
plus 2,000 rocks, 16^3 bricks) at 512^3 with the reference's default chunk
of 256 and with 320.
"""
import numpy as np
import pytest

import oracle
from tests._compare import compare_results
from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.voxelizer import voxelize_scene

pytestmark = pytest.mark.gpu


def _check(scene, chunk, chain=False):
    print(f"[chunks] {scene.name} chunk={chunk} chain={chain} T={scene.ntri}", flush=True)
    c3 = (chunk, chunk, chunk)
    ref = oracle.run(scene, flags=oracle.REF_MIP_CHAIN if chain else 0, chunk=c3)
    vz = voxelize_scene(scene, options=VOX_OPT_MIP_CHAIN if chain else 0, chunk=c3)
    try:
        assert vz.status == ref.status
        res = vz.result()
        gl = [vz.level(k) for k in range(ref.levels)]
        for k in range(ref.levels):
            assert gl[k][0].shape == ref.mips[k][0].shape, (k, gl[k][0].shape, ref.mips[k][0].shape)
        errs = compare_results(res, ref, ref_mips=ref.mips, gpu_levels=gl)
        assert not errs, "\n".join(errs[:20])
    finally:
        vz.close()
    return ref


@pytest.mark.parametrize("chunk", [8, 16])
def test_mixed_small_chunks(chunk):
    # 4 (or 2) chunks per axis; at level 1 a chunk is 4 voxels wide against
    # 6-voxel stored bricks, so bricks are assembled across three chunks
    _check(scenes.scene_mixed(seed=3, N=32, B=4), chunk)


def test_mixed_small_chunks_mip_chain():
    _check(scenes.scene_mixed(seed=3, N=32, B=4), 8, chain=True)


def test_more_groups_than_one_binning_launch_holds():
    # two launches each of setup and scatter in every chunk
    _check(scenes.scene_many_groups(seed=300), 32)


@pytest.mark.parametrize("chunk", [32, 48])
def test_mixed_chunks_uncovered_and_wide(chunk):
    # 32: the chunks reach 64 of the 80 voxels per axis; 48: they reach 96
    _check(scenes.scene_mixed(seed=5, N=64, B=8), chunk)


def test_terrain_fill_cut_at_chunk_seams():
    _check(scenes.scene_terrain_n(32, 64, 8, seed=4), 16)


def test_small_terrain_chunk_64():
    # the 128^3 / 8^3 terrain with 64^3 chunks (its default 256 covers it)
    _check(scenes.scene_small_terrain(), 64)


@pytest.mark.parametrize("chunk", [128, 256])
def test_terrain_256_b8(chunk):
    # 256^3 with 8^3 bricks: S = 320, so the reference's default chunk of 256
    # leaves the last 64 layers unvoxelized
    _check(scenes.scene_terrain_n(128, 256, 8), chunk)


def test_terrain_256_b8_mip_chain():
    # levels 1-3 filtered per chunk through the single-chunk filter with
    # chunk offsets of zero (each level reads the one below)
    _check(scenes.scene_terrain_n(128, 256, 8), 64, chain=True)


@pytest.mark.parametrize("chunk", [256, 320])
def test_full_octree_generator_512_chunks(chunk):
    ref = _check(scenes.scene_full_octree(cells=512, N=512, B=16), chunk)
    assert sum(t["nb"] for t in ref.tree) > 0
