"""Batched voxelization in one hipGraph launch (GPU): every replay of the
captured graph reproduces each job's eager result bit-exactly, with eager
work on other handles between the replays; the batched-mesh generator's jobs
(terrain tiles and random soups of uneven sizes, brick edges 8 and 16, all
voxelized by one persistent launch) equal the oracle, and equal the graphs
that voxelize each job with launches of its own, or only the small jobs in
one persistent launch; a job whose triangles
changed is refused before the graph runs."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import oracle
from tests._compare import compare_results
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


def test_generator_jobs_in_one_persistent_launch_equal_oracle():
    # brick edges 8 and 16 mixed (the oracle runs each scene at its job's)
    scs = [dataclasses.replace(sc, B=(16,) * 3 if i % 3 == 1 else (8,) * 3)
           for i, sc in enumerate(scenes.batch_scenes(seed=4, n_jobs=10, tmax=2e5,
                                                      grids=(64, 128, 256)))]
    jobs = []
    b = None
    try:
        for i, sc in enumerate(scs):
            print(f"[batch] job {i} {sc.name} T={sc.ntri} B={sc.B[0]}", flush=True)
            jobs.append(make_job(sc, sc.B[0]))
        eager = [snapshot(v) for v in jobs]
        b = Batch(jobs)
        b.capture()
        st = b.stats()
        assert st.voxelize_launches == 1 and st.voxelize_blocks > 0
        print(f"[batch] persistent: {st.graph_nodes} graph nodes", flush=True)
        for v in jobs:
            scribble(v)
        assert b.launch(), b.getErrorMessage()
        pers = [snapshot(v) for v in jobs]
        for i, (sc, v) in enumerate(zip(scs, jobs)):
            assert same(pers[i], eager[i]), f"job {i} differs from its eager run"
            ref = oracle.run(sc, keep_mips=False)
            assert ref.status == 1
            errs = compare_results(v.result(), ref)
            assert not errs, f"job {i} {sc.name}: " + "\n".join(errs[:10])
        # the same jobs voxelized by launches of their own
        b.set_voxelize(persistent=False)
        b.capture()
        assert b.stats().voxelize_launches == len(jobs)
        for v in jobs:
            scribble(v)
        assert b.launch(), b.getErrorMessage()
        for i, v in enumerate(jobs):
            assert same(snapshot(v), pers[i]), f"job {i}: per-job launches differ"
        # the small jobs in the persistent launch, the others on their own
        small = sorted(v.stats().active_tiles[0] for v in jobs)[len(jobs) // 2]
        b.set_voxelize(persistent=int(small))
        b.capture()
        n_small = sum(v.stats().active_tiles[0] <= small for v in jobs)
        assert 0 < n_small < len(jobs)
        assert b.stats().voxelize_launches == 1 + len(jobs) - n_small
        for v in jobs:
            scribble(v)
        assert b.launch(), b.getErrorMessage()
        for i, v in enumerate(jobs):
            assert same(snapshot(v), pers[i]), f"job {i}: mixed launches differ"
    finally:
        if b is not None:
            b.close()
        for v in jobs:
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
