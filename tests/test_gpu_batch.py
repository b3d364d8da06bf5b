"""Batched voxelization in one hipGraph launch (GPU): every replay of the
captured graph reproduces each job's eager result bit-exactly, with eager
work on other handles between the replays; a job whose triangles changed is
refused before the graph runs."""
import ctypes as C

import numpy as np
import pytest

from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.batch import Batch, make_job

pytestmark = pytest.mark.gpu


def snapshot(vz):
    r = vz.result()
    out = [r["occupancy"].copy()]
    for t in r["tree"]:
        out += [t["types"].copy(), t["const"].copy(), t["bidx"].copy(), t["bcol"].copy(),
                t["bnrm"].copy()]
    return out


def scribble(vz):
    """Overwrite a job's result arrays on the device, so only a replay that
    recomputes them can restore them."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    r = vz.raw_result()
    e = [b + 2 for b in r.brick_dim]
    assert hip.hipMemset(r.occupancy, 0x5A, r.occupancy_words * 4) == 0
    for d in range(r.levels):
        assert hip.hipMemset(r.types[d], 0x5A, r.nodes[d]) == 0
        assert hip.hipMemset(r.const_colors[d], 0x5A, r.nodes[d] * 4) == 0
        if r.bricks[d]:
            assert hip.hipMemset(r.brick_colors[d], 0x5A, r.bricks[d] * e[0] * e[1] * e[2] * 4) == 0
    assert hip.hipDeviceSynchronize() == 0


def same(a, b):
    return len(a) == len(b) and all(x.shape == y.shape and (x == y).all() for x, y in zip(a, b))


def test_graph_replays_match_eager():
    jobs_spec = [(scenes.scene_mixed(), 8), (scenes.scene_mixed(seed=9, N=64, B=4), 4),
                 (scenes.scene_small_terrain(), 8), (scenes.scene_two_triangles(), 8)]
    jobs = [make_job(sc, B, options=VOX_OPT_MIP_CHAIN if i % 2 else 0)
            for i, (sc, B) in enumerate(jobs_spec)]
    other = make_job(scenes.scene_mixed(seed=3), 8)
    b = None
    try:
        eager = [snapshot(v) for v in jobs]
        b = Batch(jobs)
        b.capture()
        st = b.stats()
        print(f"[batch] graph nodes {st.graph_nodes} capture {st.ms_capture:.1f} ms", flush=True)
        for rep in range(3):
            for v in jobs:
                scribble(v)
            assert not same(snapshot(jobs[0]), eager[0])
            assert b.launch(), b.getErrorMessage()
            for i, v in enumerate(jobs):
                assert same(snapshot(v), eager[i]), f"replay {rep} job {i}"
            # eager work elsewhere between replays
            other.computeEdgesFaceNormalsAndBounds()
            assert other.generateVoxelsAndOctTree(None) == 1
        assert b.stats().launches == 3
    finally:
        if b is not None:
            b.close()
        for v in jobs + [other]:
            v.close()


def test_changed_triangles_are_refused():
    sc = scenes.scene_mixed()
    vz = make_job(sc, 8)
    b = None
    try:
        b = Batch([vz])
        b.capture()
        assert b.launch()
        # same sizes, moved vertices: the checksum differs, nothing runs
        m = sc.meshes[0]
        moved = np.array(m.verts, np.float32)
        moved[0, 0, 0] += np.float32(1e-3)
        vz.deallocateTriangleMemory()
        vz.addTriangleGroup(moved, m.vnormals, m.uvs, *m.bbox(), m.is_terrain, m.image,
                            *m.addr)
        for mm in sc.meshes[1:]:
            vz.addMesh(mm)
        assert vz.allocateTriangleMemory()
        assert not b.launch()
        assert "changed" in b.getErrorMessage()
    finally:
        if b is not None:
            b.close()
        vz.close()
