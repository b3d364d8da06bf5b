"""Compressed brick output on the GPU: the device BC3 / BC1 blocks equal the
compression oracle's encoding of the device bricks byte for byte, and the
compressed .gvb files carry exactly those blocks (read back the way the
reference's LoadBinaryBricks reads them)."""
import numpy as np
import pytest

import oracle
from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.voxelizer import voxelize_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene", [scenes.scene_mixed(), scenes.scene_mixed(seed=9, N=64, B=4),
                                   scenes.scene_small_terrain()], ids=lambda s: s.name)
def test_compressed_bricks_match_oracle(scene, tmp_path):
    print(f"[bc] {scene.name}", flush=True)
    vz = voxelize_scene(scene, options=VOX_OPT_MIP_CHAIN, out_dir=str(tmp_path), compressed=True)
    try:
        res = vz.result()
        comp = vz.compressed()
        assert comp is not None and len(comp) == res["levels"]
        e = tuple(b + 2 for b in scene.B)
        ref = oracle.pool_load(str(tmp_path), slots=(4, 4, 4))
        assert ref.compressed
        total = 0
        for d, tr in enumerate(res["tree"]):
            oc, on = oracle.bc_encode_bricks(tr["bcol"], tr["bnrm"], e)
            gc, gn = comp[d]
            assert (gc == oc).all() and (gn == on).all(), f"depth {d}"
            fc, fn = ref.file_compressed[d]
            assert (fc == gc).all() and (fn == gn).all(), f"file depth {d}"
            total += tr["nb"]
        assert total > 0
    finally:
        vz.close()
