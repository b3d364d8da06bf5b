"""Asynchronous generate (GPU): several handles voxelizing on worker threads
at once, with eager work on the calling thread meanwhile, give each handle
the result it gets alone, bit-exactly; releasing the work memory keeps the
result."""
import numpy as np
import pytest

from voxvizthesis_amd import scenes
from voxvizthesis_amd.voxelizer import Voxelizer

pytestmark = pytest.mark.gpu


def upload(sc, B):
    vz = Voxelizer(sc.N, (B, B, B))
    assert vz.initCuda() and vz.allocateVoxelMemory()
    assert vz.setVoxelizationParams(sc.p, sc.delta)
    for m in sc.meshes:
        assert vz.addMesh(m)
    assert vz.allocateTriangleMemory()
    return vz


def snapshot(vz):
    r = vz.result()
    out = [r["occupancy"]]
    for t in r["tree"]:
        out += [t["types"], t["const"], t["bidx"], t["bcol"], t["bnrm"]]
    return out


def same(a, b):
    return len(a) == len(b) and all(x.shape == y.shape and (x == y).all() for x, y in zip(a, b))


def test_async_matches_alone():
    specs = [(scenes.scene_mixed(), 8), (scenes.scene_small_terrain(), 8),
             (scenes.scene_mixed(seed=9, N=64, B=4), 4)]
    alone = []
    for sc, B in specs:
        vz = upload(sc, B)
        try:
            assert vz.generateVoxelsAndOctTree(None) == 1, vz.getErrorMessage()
            alone.append(snapshot(vz))
        finally:
            vz.close()
    print("[async] eager references done", flush=True)
    jobs = [upload(sc, B) for sc, B in specs]
    other = upload(scenes.scene_two_triangles(), 8)
    try:
        for vz in jobs:
            assert vz.generateAsync(), vz.getErrorMessage()
        # the calling thread keeps the device busy with another handle
        assert other.generateVoxelsAndOctTree(None) == 1
        for i, vz in enumerate(jobs):
            st = vz.wait(True)
            assert st == 1, (i, vz.getErrorMessage())
            assert same(snapshot(vz), alone[i]), f"job {i}"
        print("[async] concurrent results match", flush=True)
        # a second round on the same handles, then the work memory is freed
        for vz in jobs:
            assert vz.computeEdgesFaceNormalsAndBounds()
            assert vz.generateAsync()
        for i, vz in enumerate(jobs):
            assert vz.wait(True) == 1
            assert vz.releaseWorkMemory(), vz.getErrorMessage()
            assert same(snapshot(vz), alone[i]), f"job {i} after release"
        assert jobs[0].wait(True) == 0  # nothing running any more
    finally:
        for vz in jobs + [other]:
            vz.close()


def test_busy_handle_refuses_calls():
    sc = scenes.scene_mixed()
    vz = upload(sc, 8)
    try:
        assert vz.generateAsync()
        st = vz.wait(False)
        if st == 2:
            # still running: other calls fail without touching the handle
            assert not vz.setVoxelizationParams(sc.p, sc.delta)
            assert "vox_wait" in vz.getErrorMessage()
        assert vz.wait(True) == 1
        assert vz.setVoxelizationParams(sc.p, sc.delta)
        r = vz.result()
        assert int(np.unpackbits(r["occupancy"].view(np.uint8)).sum()) > 0
    finally:
        vz.close()


def test_timing_run_equals_overlapped_run():
    """A generate queues its occupancy bitmap and collapse on a second stream
    beside the mip chain and the brick gather; a timing run keeps them on the
    handle's stream.  Both give the same result, run after run on one
    handle, including the occupancy bitmap and the collapsed types."""
    from voxvizthesis_amd._lib import VOX_OPT_TIMING
    for sc, B in ((scenes.scene_small_terrain(), 8), (scenes.scene_mixed(seed=9, N=64, B=4), 4)):
        print(f"[async] timing vs overlapped {sc.name}", flush=True)
        vz = upload(sc, B)
        try:
            assert vz.generateVoxelsAndOctTree(None) == 1, vz.getErrorMessage()
            ref = snapshot(vz)
            vz._L.vox_set_options(vz._h, VOX_OPT_TIMING)
            assert vz.computeEdgesFaceNormalsAndBounds()
            assert vz.generateVoxelsAndOctTree(None) == 1, vz.getErrorMessage()
            assert same(snapshot(vz), ref)
            vz._L.vox_set_options(vz._h, 0)
            for _ in range(2):
                assert vz.computeEdgesFaceNormalsAndBounds()
                assert vz.generateVoxelsAndOctTree(None) == 1, vz.getErrorMessage()
                assert same(snapshot(vz), ref)
            assert any(t.any() for t in ref[1::5])   # some node typed
        finally:
            vz.close()
