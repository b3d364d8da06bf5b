"""Many voxelizations in one hipGraph launch, and the autotuner that picks
each job's brick edge and persistent launch shape (include/voxmi_batch.h).

A job is a Voxelizer that has run one eager generateVoxelsAndOctTree (which
sizes its buffers).  `Batch.capture()` records every job's triangle setup and
generate into a single graph, with the voxelization of all jobs in one
persistent launch over a device work queue; `Batch.launch()` replays it after
checking on the device that no job's triangles changed.
"""
from __future__ import annotations

import ctypes as C
import math
import time

from . import _lib
from .voxelizer import Voxelizer


class Batch:
    def __init__(self, voxelizers):
        self._L = _lib.load()
        self.jobs = list(voxelizers)
        arr = (C.c_void_p * len(self.jobs))(*[v._h for v in self.jobs])
        self._h = self._L.vox_batch_create(arr, len(self.jobs))
        if not self._h:
            raise MemoryError("vox_batch_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    def getErrorMessage(self) -> str:
        return self._L.vox_batch_get_error(self._h).decode()

    def set_voxelize(self, persistent=-1, blocks: int = 0):
        """Which jobs share one persistent voxelize launch per wave of groups
        (blocks 0 = 4 workgroups per CU): -1 / True every job, 0 / False
        none (a launch per job and group), n > 0 the jobs with at most n
        active level-0 tiles."""
        if persistent is True:
            persistent = -1
        if not self._L.vox_batch_set_voxelize(self._h, int(persistent), int(blocks)):
            raise RuntimeError(self.getErrorMessage())

    def capture(self):
        if not self._L.vox_batch_capture(self._h):
            raise RuntimeError(self.getErrorMessage())

    def launch(self) -> bool:
        """True when the graph ran; False (with getErrorMessage) when a job's
        inputs changed since the capture."""
        return bool(self._L.vox_batch_launch(self._h))

    def stats(self) -> _lib.VoxBatchStats:
        s = _lib.VoxBatchStats()
        self._L.vox_batch_get_stats(self._h, C.byref(s))
        return s

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def brick_candidates(N: int, edges=(8, 16, 32)):
    """Brick edges B for which N / B is a power of two >= 2."""
    out = []
    for b in edges:
        if N % b == 0:
            r = N // b
            if r >= 2 and (r & (r - 1)) == 0:
                out.append(b)
    return out


def make_job(scene, B: int, persistent_blocks: int = 0, options: int = 0) -> Voxelizer:
    """Upload a scene into a voxelizer with brick edge B and run the eager
    generate that sizes it."""
    vz = Voxelizer(scene.N, (B, B, B), None, options=options)
    ok = vz.initCuda() and vz.allocateVoxelMemory() and vz.setVoxelizationParams(scene.p,
                                                                              scene.delta)
    for m in scene.meshes:
        ok = ok and vz.addMesh(m)
    ok = ok and vz.allocateTriangleMemory()
    if persistent_blocks:
        ok = ok and bool(vz._L.vox_set_launch(vz._h, int(persistent_blocks)))
    ok = ok and vz.computeEdgesFaceNormalsAndBounds()
    if not ok or vz.generateVoxelsAndOctTree(None) == 0:
        err = vz.getErrorMessage()
        vz.close()
        raise RuntimeError(err)
    return vz


def autotune(scene, edges=(8, 16, 32), blocks_per_cu=(4, 8, 16), cus: int = 256, reps: int = 2,
             keep: bool = False):
    """Time eager generates of the scene for every brick edge and persistent
    launch shape (one voxelizer per brick edge, its launch shape changed
    between timings); return (B, persistent_blocks, ms) of the fastest, and
    with keep=True also that voxelizer, set to the chosen shape (the caller
    closes it)."""
    best, best_vz = None, None
    try:
        for B in brick_candidates(scene.N[0], edges):
            vz = make_job(scene, B)
            won = False
            try:
                for k in blocks_per_cu:
                    pb = k * cus
                    if not vz._L.vox_set_launch(vz._h, int(pb)):
                        raise RuntimeError(vz.getErrorMessage())
                    t = math.inf
                    for _ in range(reps):
                        vz.computeEdgesFaceNormalsAndBounds()
                        t0 = time.perf_counter()
                        if vz.generateVoxelsAndOctTree(None) == 0:
                            raise RuntimeError(vz.getErrorMessage())
                        t = min(t, (time.perf_counter() - t0) * 1e3)
                    if best is None or t < best[2]:
                        best = (B, pb, t)
                        won = True
            except BaseException:
                vz.close()
                raise
            if keep and won:
                if best_vz is not None:
                    best_vz.close()
                best_vz = vz
            else:
                vz.close()
    except BaseException:
        if best_vz is not None:
            best_vz.close()
        raise
    if keep:
        if best_vz is not None:
            # the kept voxelizer's last timing may have used another shape
            best_vz._L.vox_set_launch(best_vz._h, int(best[1]))
            best_vz.computeEdgesFaceNormalsAndBounds()
            if best_vz.generateVoxelsAndOctTree(None) == 0:
                err = best_vz.getErrorMessage()
                best_vz.close()
                raise RuntimeError(err)
        return best[0], best[1], best[2], best_vz
    return best
