"""The full-octree workload at its real size on the GPU (1025^2 fBm terrain +
2,000 rocks, 2.26 M triangles, 1024^3 grid, 16^3 bricks, depths 0-6), checked
through properties that do not depend on size (the oracle needs ~30 GB of
dense arrays and minutes of CPU at this size; it is compared at small sizes
in test_gpu_parity.py):

* order independence: two voxelizations with different persistent grids
  (so different tile-to-workgroup assignments, hit-record segments and
  atomic orders) give bit-identical occupancy, node arrays and brick
  payloads;
* brick addressing: the NON-CONST nodes of a depth, in node order, hold
  brick indices 0..nb-1; every other node holds none;
* CollapseConstantBranches (Voxelizer.cu 701-774): a node is LEAF-CONST (2)
  exactly when it is CONST and all eight children are constant (CONST
  leaves, LEAF-CONST inner nodes);
* a NON-CONST node's brick is not one colour (its typing found a voxel
  different from the node's sample)."""
import numpy as np
import pytest

from voxvizthesis_amd import scenes
from voxvizthesis_amd.voxelizer import Voxelizer

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def run(scene, blocks):
    vz = Voxelizer(scene.N, scene.B, None)
    ok = vz.initCuda() and vz.allocateVoxelMemory() and vz.setVoxelizationParams(scene.p,
                                                                              scene.delta)
    for m in scene.meshes:
        ok = ok and vz.addMesh(m)
    ok = ok and vz.allocateTriangleMemory()
    assert ok, vz.getErrorMessage()
    assert vz._L.vox_set_launch(vz._h, blocks)
    assert vz.computeEdgesFaceNormalsAndBounds()
    assert vz.generateVoxelsAndOctTree(None) == 1, vz.getErrorMessage()
    return vz


def test_full_octree_order_independent_and_consistent():
    sc = scenes.scene_full_octree()
    print(f"[fullsize] {sc.name}: {sc.ntri} triangles", flush=True)
    va = run(sc, 2048)
    try:
        a = va.result()
    finally:
        va.close()
    vb = run(sc, 333)
    try:
        b = vb.result()
    finally:
        vb.close()
    print(f"[fullsize] bricks per depth {[t['nb'] for t in a['tree']]}", flush=True)
    assert (a["occupancy"] == b["occupancy"]).all()
    assert int(np.unpackbits(a["occupancy"].view(np.uint8)).sum()) > 0
    L = a["levels"]
    assert L == 7
    for d in range(L):
        ta, tb = a["tree"][d], b["tree"][d]
        assert ta["nb"] == tb["nb"], f"depth {d}"
        for k in ("types", "const", "bidx", "bcol", "bnrm"):
            assert (ta[k] == tb[k]).all(), f"depth {d} {k}"
    assert sum(t["nb"] for t in a["tree"]) > 10000
    for d in range(L):
        t = a["tree"][d]
        types, bidx, nb = t["types"], t["bidx"], t["nb"]
        non = types == 1
        assert int(non.sum()) == nb, f"depth {d}"
        assert (bidx[non] == np.arange(nb, dtype=np.uint32)).all(), f"depth {d}"
        assert (bidx[~non] == NONE).all(), f"depth {d}"
        if d == L - 1:
            assert set(np.unique(types)) <= {0, 1}
        else:
            Rd = 1 << d
            child = a["tree"][d + 1]["types"].reshape(Rd, 2, Rd, 2, Rd, 2)
            want = 0 if d + 1 == L - 1 else 2
            all_const = (child == want).all(axis=(1, 3, 5)).reshape(-1)
            assert ((types == 2) == ((types != 1) & all_const)).all(), f"depth {d}"
        if nb:
            bc = t["bcol"]
            assert (bc.min(axis=1) != bc.max(axis=1)).all(), f"depth {d}: a one-colour brick"
