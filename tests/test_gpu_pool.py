"""Device node pool / brick pool builder against the pool oracle, bit-exact
(GPU).  The oracle reads the files the voxelizer wrote (tree.gvx, .gvb) the
way the reference's reader does; the device builder works from the result in
HBM.  Equality therefore also pins the file writer against the in-memory
result."""
import numpy as np
import pytest

import oracle
from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.gvpool import BrickPool, OctTree
from voxvizthesis_amd.voxelizer import voxelize_scene

pytestmark = pytest.mark.gpu


def _case(scene, tmp_path, slots, mode="text", options=0):
    print(f"[pool] {scene.name} slots={slots} mode={mode} options={options}", flush=True)
    vz = voxelize_scene(scene, options=options)
    bp = t = None
    try:
        assert vz.status == 1
        assert vz.writeFiles(str(tmp_path))
        ref = oracle.pool_load(str(tmp_path), slots=slots, store_const=(mode != "reference"))
        bp = BrickPool(scene.B, slots)
        t = OctTree(bp, vz, const_mode=mode)
        assert t.get3DTextureDimensions() == ref.node_dims
        assert t.getNodeCount() == 8 * ref.blocks
        nodes = t.nodes()
        bad = np.argwhere((nodes != ref.nodes).any(axis=-1))
        assert bad.size == 0, f"{len(bad)} texels differ, first {bad[:4].tolist()}"
        col, nrm = bp.textures()
        assert (col == ref.brick_colors).all() and (nrm == ref.brick_normals).all()
        res = vz.result()
        for d, tr in enumerate(res["tree"]):
            fc, fn = ref.file_bricks[d]
            assert fc.shape[0] == tr["nb"]
            if tr["nb"]:
                assert (fc == tr["bcol"]).all() and (fn == tr["bnrm"]).all()
            # node_texel names the texel of every node in the tree
            txl = t.node_texel(d)
            inside = txl != 0xFFFFFFFF
            w0 = nodes.reshape(-1, 2)[txl[inside], 0]
            assert ((w0 & 0x40000000) != 0).tolist() == (tr["types"][inside] != 1).tolist()
        return vz, t, bp, res, nodes
    except Exception:
        for o in (t, bp, vz):
            if o is not None:
                o.close()
        raise


def test_small_terrain_pool(tmp_path):
    vz, t, bp, res, _ = _case(scenes.scene_small_terrain(), tmp_path, (16, 16, 16))
    for o in (t, bp, vz):
        o.close()


def test_small_terrain_chain_root_brick(tmp_path):
    # with the mip chain the root holds a brick: it goes to slot 0
    vz, t, bp, res, nodes = _case(scenes.scene_small_terrain(), tmp_path, (4, 4, 4),
                                  options=VOX_OPT_MIP_CHAIN)
    try:
        assert res["tree"][0]["types"][0] == 1
        assert nodes[0, 0, 0, 0] & 0x80000000 and nodes[0, 0, 0, 1] == (1 << 20) | (1 << 10) | 1
        assert bp.view().resident == 1
    finally:
        for o in (t, bp, vz):
            o.close()


def test_mixed_reference_const(tmp_path):
    vz, t, bp, res, nodes = _case(scenes.scene_mixed(), tmp_path, (8, 8, 8), mode="reference",
                                  options=VOX_OPT_MIP_CHAIN)
    try:
        # constant nodes carry no colour, as the reference's reader leaves them
        w = nodes.reshape(-1, 2)
        const = (w[:, 0] & 0x40000000) != 0
        assert not w[const, 1].any()
    finally:
        for o in (t, bp, vz):
            o.close()


def test_mixed_b4_exact_const(tmp_path):
    sc = scenes.scene_mixed(seed=9, N=64, B=4)
    print(f"[pool] {sc.name} exact", flush=True)
    vz = voxelize_scene(sc, options=VOX_OPT_MIP_CHAIN)
    bp = BrickPool(sc.B, (8, 8, 8))
    t = OctTree(bp, vz, const_mode="exact")
    try:
        res = vz.result(with_bricks=False)
        nodes = t.nodes().reshape(-1, 2)
        for d, tr in enumerate(res["tree"]):
            txl = t.node_texel(d)
            sel = (txl != 0xFFFFFFFF) & (tr["types"] != 1)
            assert (nodes[txl[sel], 1] == tr["const"][sel]).all()
    finally:
        for o in (t, bp, vz):
            o.close()
