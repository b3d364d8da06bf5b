"""HIP path against the oracle, bit-exact (GPU)."""
import numpy as np
import pytest

import oracle
from tests._compare import compare_results
from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.voxelizer import voxelize_scene

pytestmark = pytest.mark.gpu


def _check(scene, chain=False, levels=True):
    print(f"[parity] {scene.name} chain={chain} T={scene.ntri}", flush=True)
    ref = oracle.run(scene, flags=oracle.REF_MIP_CHAIN if chain else 0, keep_mips=levels)
    vz = voxelize_scene(scene, options=VOX_OPT_MIP_CHAIN if chain else 0)
    try:
        assert vz.status == ref.status
        res = vz.result()
        gl = [vz.level(k) for k in range(ref.levels)] if levels else None
        errs = compare_results(res, ref, ref_mips=ref.mips, gpu_levels=gl)
        assert not errs, "\n".join(errs[:20])
    finally:
        vz.close()
    return ref


def test_two_triangles():
    ref = _check(scenes.scene_two_triangles())
    assert int(np.unpackbits(ref.occupancy.view(np.uint8)).sum()) > 0


def test_small_terrain():
    _check(scenes.scene_small_terrain())


def test_small_terrain_mip_chain():
    _check(scenes.scene_small_terrain(), chain=True)


def test_mixed_groups():
    _check(scenes.scene_mixed())


def test_mixed_groups_b4():
    _check(scenes.scene_mixed(seed=9, N=64, B=4))


def test_mixed_groups_rocks_first():
    # the first group is no terrain: every level-0 voxel no group reaches is zero
    _check(scenes.scene_mixed(seed=5, order="rocks_first"))


def test_mixed_groups_two_terrains():
    # a second terrain fills over voxels that the first fill and the rocks hold
    _check(scenes.scene_mixed(seed=6, order="two_terrains"))


def test_more_groups_than_one_binning_launch_holds():
    # setup and scatter take at most 8 groups per launch: 11 groups with
    # triangles (10 inside the grid) need two launches of each, and the
    # terrain's fill runs in the second
    _check(scenes.scene_many_groups())


def test_big_and_small_triangles_binned():
    # boxes of 18^3 tiles walk their bins; the small ones are listed as pairs
    _check(scenes.scene_big_triangles())


def test_type_x_split_miss():
    # N=256, B=4: the leaf typing launch split leaves x = N + extra (the +1
    # border of the last node column) unexamined.
    _check(scenes.scene_terrain_n(128, 256, 4), levels=False)


def test_brick_sizes_not_powers_of_two():
    # (B + 2)^3 odd (B = 3, 5): the brick gather writes single words; B = 6
    # writes 16-byte runs of a region that starts mid-tile
    for N, B in ((48, 3), (80, 5), (96, 6)):
        _check(scenes.scene_terrain_n(64, N, B))


def test_full_octree_generator_256():
    # the headline workload's own generator (fBm terrain + 2,000 rocks, 16^3
    # bricks) at a size the oracle runs in seconds: 256^3, depths 0-4
    _check(scenes.scene_full_octree(cells=256, N=256, B=16))


def test_full_octree_generator_512():
    # the same at 512^3 (depths 0-5, 526 K triangles), every level compared
    _check(scenes.scene_full_octree(cells=512, N=512, B=16))


def test_reference_default_brick_edge_30():
    # main.cu's default brick edge of 30 (32^3 stored bricks), N / B = 8
    _check(scenes.scene_terrain_n(120, 240, 30))


def test_brick_wider_than_the_gather_tile_table():
    # B = 64: a stored brick of 66^3 spans 9 tiles per axis, more than the
    # brick gather stages in LDS (6), so it reads the tile map directly
    _check(scenes.scene_terrain_n(64, 128, 64))
