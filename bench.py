#!/usr/bin/env python3
"""Voxelizer benchmark: one JSON line on stdout.

Workload (one GPU): the full-octree terrain scene -- 1025x1025 fBm heightfield
(2,097,152 textured triangles) plus 2,000 icosphere rocks (160,000 untextured
triangles) voxelized into a 1024^3 grid of 16^3 bricks (18^3 stored), octree
depths 0-6, constant bricks culled.  A step is one pass of the hot path with
the triangles already resident in HBM: triangle setup and binning
(computeEdgesFaceNormalsAndBounds) followed by voxelization, colour, fill,
mip chain, node typing, brick gather and collapse (generateVoxelsAndOctTree),
ending with the bricks, node arrays and occupancy bitmap in HBM.

value = N^3 voxels classified per second (Gvoxels/s).  The roofline object
describes k_voxelize, the kernel with the largest share of device time in the
rocprof trace of this workload, timed with HIP events on the library's
stream; "rooflines" gives the other kernels the same way (brick gather, mip
filter, and the binning passes: setup, scan, scatter).  Algorithmic bytes per
launch are given by kernel_bytes() below.  cpu_baseline times the oracle (a
scalar C++ port of the reference logic) on the same workload, tiled into
256^3 windows at its own voxel size and spread over every host core the
process may use (at most 16).

Further legs (each skippable): serving the octree to ray-cast camera streams
(a pool that holds every brick, and one smaller, so the LRU evicts), paged
serving over an 8192^3 and a 32768^3 virtual grid with tiles built and
voxelized on demand, and batched voxelization of 256 meshes of uneven sizes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gvoxels/s + bricks/s voxelized (HBM GB/s % peak); p99 brick-request latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(scene, st, res_bricks, B):
    """Compulsory HBM bytes of one step (the pipeline formula):
    96*T + 16*P + 8*leaf_nodes + N^3/8 + sum_k bricks_k*(B+2)^3*8
    + sum_{k>=1} bricks_{k-1}*(B+2)^3*8 + 8*nodes."""
    T = st.triangles
    P = st.pairs
    N = scene.N[0]
    L = len(res_bricks)
    bvox = (B + 2) ** 3
    leaf_nodes = 8 ** (L - 1)
    nodes = sum(8 ** d for d in range(L))
    bricks = sum(res_bricks)
    child_reads = sum(res_bricks[d + 1] for d in range(L - 1))
    return (96 * T + 16 * P + 8 * leaf_nodes + N ** 3 // 8 + bricks * bvox * 8
            + child_reads * bvox * 8 + 8 * nodes)


def kernel_bytes(name, scene, st, bricks, B, launches):
    """Algorithmic bytes per launch of the named kernel."""
    bvox = (B + 2) ** 3
    if name == "brick_gather":
        # read (B+2)^3 tile voxels (RGBA8 colour + packed RGB8 normal), write
        # the brick's RGBA8 colour + RGB8(+pad) normal
        return sum(bricks) * bvox * (8 + 8) / launches
    if name == "voxelize":
        # triangle attributes read once (vertices 36 B, vertex normals 36 B,
        # UVs 24 B where the group has them) + (triangle, tile) bin entries
        # written and read (16 B each, the pipeline formula's 16*P)
        tri_bytes = 0
        for m in scene.meshes:
            tri_bytes += m.ntri * (36 + (36 if m.vnormals is not None else 0)
                                   + (24 if m.uvs is not None else 0))
        return (tri_bytes + 16 * st.pairs) / launches
    if name == "mip":
        # a filtered level reads the stored voxels of the active source tiles
        # under its active tiles (RGBA8 colour + fp32 normal, 16 B; level 0
        # keeps only hit or filled voxels, counted on the device in the timing
        # step) and writes colour, packed and fp32 normal (20 B) for each voxel
        # of its active tiles.  Without the mip-chain option the even levels
        # >= 2 alias the level below them and are not filtered.
        a = st.active_tiles
        tot = 0
        for k in range(1, len(bricks)):
            if k >= 2 and k % 2 == 0:
                continue
            tot += st.mip_source_voxels[k] * 16 + a[k] * 512 * 20
        return tot / launches
    if name == "setup":
        # vertices read (36 B / triangle), face normal + packed voxel bounds
        # written (24 B), and the count of every (group, tile) bin that a
        # triangle reaches written (4 B; at most one per (triangle, tile)
        # pair, the bound used here).  Empty bins are never touched.
        return (60 * st.triangles + 4 * st.pairs) / launches
    if name == "scan":
        # every prefix sum reads and writes each value (8 B); the bin scan
        # also zeroes its input (4 B)
        return (8 * st.scan_values + 4 * st.bins) / launches
    if name == "scatter":
        # packed bounds read (12 B / triangle) and one triangle index written
        # per (triangle, tile) bin entry (4 B); the bins' start and fill words
        # it also reads and updates are not counted (a few per tile, mostly
        # in L2)
        return (12 * st.triangles + 4 * st.pairs) / launches
    return 0.0


def copy_reference(nbytes: int, reps: int):
    """GB/s of a plain device copy of nbytes (read + write counted), timed
    on the host around each synchronised copy; the median of reps."""
    import numpy as np
    from voxvizthesis_amd import _lib
    L = _lib.load()
    src, dst = L.vox_device_alloc(nbytes), L.vox_device_alloc(nbytes)
    try:
        if not (src and dst):
            raise RuntimeError("copy reference: device allocation failed")
        ts = []
        for i in range(reps + 1):
            t0 = time.perf_counter()
            if not L.vox_device_copy(dst, src, nbytes):
                raise RuntimeError("copy reference failed")
            if i:
                ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        gbs = 2 * nbytes / t / 1e9
        return {"bytes_per_copy": 2 * nbytes, "ms": round(t * 1e3, 4), "gbs": round(gbs, 1),
                "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
    finally:
        if src:
            L.vox_device_free(src)
        if dst:
            L.vox_device_free(dst)


def chunked_bench(scene, chunk: int, steps: int, warmup: int):
    """The headline workload with voxChunkDim = chunk^3, the reference's
    default (Voxelizer.h 124; its caller keeps it, main.cu 4169-4170): per
    chunk mips and node samples, types OR-ed over chunks, bricks across
    chunks assembled, bricks numbered in the writer's order.  Same timing as
    the headline (setup + generate per step, inputs resident)."""
    from voxvizthesis_amd.voxelizer import Voxelizer
    vz = Voxelizer(scene.N, scene.B, (chunk,) * 3)
    try:
        ok = vz.initCuda() and vz.allocateVoxelMemory() and \
            vz.setVoxelizationParams(scene.p, scene.delta)
        for m in scene.meshes:
            ok = ok and vz.addMesh(m)
        ok = ok and vz.allocateTriangleMemory()
        if not ok:
            raise RuntimeError(vz.getErrorMessage())

        def step():
            if not vz.computeEdgesFaceNormalsAndBounds() or \
                    vz.generateVoxelsAndOctTree(None) != 1:
                raise RuntimeError(vz.getErrorMessage())

        for _ in range(warmup):
            step()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ms = (time.perf_counter() - t0) / steps * 1e3
        res = vz.result(with_bricks=False)
        bricks = [t["nb"] for t in res["tree"]]
        n = scene.N[0] * scene.N[1] * scene.N[2]
        return {
            "workload": f"the headline workload in {chunk}^3 chunks "
                        f"({(scene.N[0] // chunk) ** 3} chunks)",
            "chunk": chunk,
            "ms_per_step": round(ms, 4),
            "gvoxels_per_s": round(n / (ms * 1e-3) / 1e9, 2),
            "bricks": bricks,
            "bricks_per_s": round(sum(bricks) / (ms * 1e-3), 1),
        }
    finally:
        vz.close()


def serving_bench(vz, scene, frames: int, W: int, H: int, slots, max_up: int, streams: int = 1):
    """Paged flythrough over the voxelized octree: per frame each camera
    stream's W x H sparse-octree ray cast emits brick requests and one brick
    server admits them all into the shared pool (LRU).  Frames are queued
    without host waits: the ray casts of a frame overlap on the device and
    the server frame waits for them.  Returns throughput and request latency
    figures."""
    import numpy as np
    from voxvizthesis_amd import scenes
    from voxvizthesis_amd.gvpool import BrickPool, OctTree
    from voxvizthesis_amd.gvserve import BrickServer, RayCaster, camera, raycast_params

    bp = BrickPool(scene.B, slots)
    t = OctTree(bp, vz)
    rcs = [RayCaster(t, W, H) for _ in range(streams)]
    srv = BrickServer(t, max_up)
    lo = np.asarray(scene.p, np.float32)
    hi = lo + np.asarray(scene.N, np.float32) * np.asarray(scene.delta, np.float32)
    prm = raycast_params(lo, hi, jitter_seed=5)
    paths = [scenes.flythrough(seed=5 + k, frames=frames) for k in range(streams)]
    ray_ms, dedup_ms, req = [], [], 0
    t0 = time.perf_counter()
    for f in range(frames):
        for rc, path in zip(rcs, paths):
            pos, look, up = path[f]
            rc.frame(camera(pos, look, up, 60.0, W, H, f), prm, wait=False)
        srv.frame(rcs, wait=False)
        for rc in rcs:  # waits for the ray casts only
            v = rc.view()
            ray_ms.append(v.ms_raycast)
            dedup_ms.append(v.ms_requests)
            req += v.nrequests
    st = srv.stats()
    wall = time.perf_counter() - t0
    lat = srv.latencies_us() / 1e3
    out = {
        "workload": f"{frames}-frame seeded Catmull-Rom flythrough (12 points, altitude "
                    f"0.05-0.2 extent) over the full octree, {streams} camera stream(s) of "
                    f"{W}x{H} ray cast",
        "frames": frames,
        "streams": streams,
        "pool_slots": int(np.prod(slots)),
        "max_uploads_per_frame": max_up,
        "requests": int(req),
        "uploads": int(st.uploads),
        "evictions": int(st.evictions),
        "pending_at_end": int(st.pending),
        "bricks_per_s": round(st.uploads / wall, 1),
        "frame_ms": round(wall / frames * 1e3, 3),
        "raycast_ms_mean": round(float(np.mean(ray_ms)), 4),
        "dedup_ms_mean": round(float(np.mean(dedup_ms)), 4),
        "update_ms_mean": round(st.ms_update_total / max(st.frames, 1), 4),
        "latency_p50_ms": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
        "latency_p99_ms": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
        "latency_max_ms": round(float(lat.max()), 4) if len(lat) else None,
    }
    for o in [srv] + rcs + [t, bp]:
        o.close()
    return out


def run_paged(pg, paths, frames: int, W: int, H: int):
    """Frames at 60 Hz over the camera paths (one per stream): queue the
    frame, wait for its server step, sleep to the next frame time.  Returns
    the per-frame wall times, the total and the loaded-tile counts."""
    import numpy as np
    period = 1.0 / 60.0
    wall, loaded_at = [], []
    t_start = time.perf_counter()
    for f in range(frames):
        cams = [dict(position=p[f][0], look=p[f][1], up=p[f][2], fov_y_deg=60.0, width=W,
                     height=H) for p in paths]
        a = time.perf_counter()
        pg.frame(cams)
        pg.server.stats()
        wall.append(time.perf_counter() - a)
        loaded_at.append(pg.tiles_loaded)
        rest = t_start + (f + 1) * period - time.perf_counter()
        if rest > 0:
            time.sleep(rest)
        if (f + 1) % 100 == 0:
            log(f"[bench] paged frame {f + 1}: {pg.tiles_loaded} tiles voxelized")
    return np.asarray(wall), time.perf_counter() - t_start, loaded_at


def paged_stats(pg, wall, total, loaded_at, period=1.0 / 60.0):
    import numpy as np
    st = pg.server.stats()
    lat = pg.server.latencies_us() / 1e3
    loads = np.asarray(pg.load_ms) if pg.load_ms else np.zeros(1)
    first = next((f for f, n in enumerate(loaded_at) if n > 0), None)
    phases = getattr(pg, "load_phases", [])
    return {
        "tiles": len(pg.trees),
        "tiles_voxelized": int(pg.tiles_loaded),
        "frame_period_ms": round(period * 1e3, 3),
        "frame_ms": round(float(np.mean(wall)) * 1e3, 3),
        "frame_ms_p99": round(float(np.percentile(wall * 1e3, 99)), 3),
        "frames_over_period": int((wall > period).sum()),
        "first_tile_frame": first,
        # a request that reaches a tile not yet in memory waits for the tile:
        # from the frame that queued its voxelization to the frame that
        # attached its octree (its root brick then goes through the LRU
        # like any other request)
        "tile_miss_ms_p50": round(float(np.percentile(loads, 50)), 2),
        "tile_miss_ms_p99": round(float(np.percentile(loads, 99)), 2),
        "tile_miss_ms_max": round(float(loads.max()), 2),
        # host time of a tile load's phases on its loader thread (median):
        # voxelizer creation and voxel memory, the tile's triangles built,
        # triangle memory + setup, generate, release of the work memory
        "tile_load_phase_ms_p50": {k: round(float(np.median([p[k] for p in phases])), 2)
                                   for k in (phases[0] if phases else {})},
        "uploads": int(st.uploads),
        "evictions": int(st.evictions),
        "bricks_per_s": round(st.uploads / total, 1),
        "latency_p50_ms": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
        "latency_p99_ms": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
    }


def wide_slots(B: int, gb: float, normals: bool = False) -> int:
    """Slots of a wide pool of gb GB of (B+2)^3 RGBA8 colour bricks (+ RGB8
    normals when `normals`)."""
    return int(gb * 1e9 // ((B + 2) ** 3 * (8 if normals else 4)))


def paged_bench(cells: int, frames: int, W: int, H: int, tile: int, B: int, pool_gb: float,
                max_trees: int, max_loading: int, serve_group: int, f2b_stop=None):
    """Served paging: a (cells+1)^2 fBm terrain in a cells^3 virtual grid,
    tiled into tile^3 octrees built and voxelized on demand on the device as
    the seeded flythrough first sees them (the pager), served from one shared
    brick pool of pool_gb GB of RGBA8 colour bricks (a wide pool) with LRU
    eviction;
    a server frame follows every serve_group trees' ray casts within a
    frame.  The terrain's heights are the scene generator's fBm (seed 5),
    built on the device rather than as a host triangle soup."""
    import numpy as np
    from voxvizthesis_amd import scenes
    from voxvizthesis_amd.terrain import TerrainPager

    # In metres with 1 m voxels, as the reference's tiler assumes (it floors
    # each tree's z origin to a whole unit, main.cu 4258); the camera planes
    # are the reference's 1 m / 10 km.
    extent = float(cells)
    slots = wide_slots(B, pool_gb)
    t0 = time.perf_counter()
    pg = TerrainPager(cells, tile, B, slots, seed=5, ridged_amp=0.0, rocks=0,
                      max_trees=max_trees, max_loading=max_loading, streams=1, max_width=W,
                      max_height=H, near=1.0, far=10000.0, serve_group=serve_group,
                      f2b_stop=f2b_stop, pool_normals=False)
    t_setup = time.perf_counter() - t0
    log(f"[bench] paged: {len(pg.trees)} paged trees set up in {t_setup:.1f} s")
    path = scenes.flythrough(seed=5, frames=frames, extent=extent, height_fn=pg.height_fn)
    try:
        wall, total, loaded_at = run_paged(pg, [path], frames, W, H)
        out = {
            "workload": f"{cells + 1}^2 fBm terrain (seed 5, {cells} m, 1 m voxels) in a "
                        f"{cells}^3 virtual grid, "
                        f"{tile}^3 tiles of {B}^3 bricks built and voxelized on demand, "
                        f"{frames}-frame seeded flythrough, {W}x{H} ray cast",
            "serve_group": serve_group,
            "cast": cast_mode(f2b_stop),
            "pool_slots": int(np.prod(slots)),
            "pool_gb": round(float(np.prod(slots)) * (B + 2) ** 3 * 4 / 1e9, 2),
            "pool_format": "wide (slot in word 1), RGBA8 colour only",
            "scene_setup_s": round(t_setup, 2),
            "frames": frames,
        }
        out.update(paged_stats(pg, wall, total, loaded_at))
        return out
    finally:
        pg.close()


def cast_mode(f2b_stop):
    if f2b_stop is None:
        return ("every tree cast alone over its screen rectangle, blended far to near (the "
                "reference's renderer); a server frame after every serve_group casts")
    return ("the visible trees cast front to back in passes of serve_group, serve_group, then "
            "doubling numbers of trees, each resuming the frame where the nearer trees left "
            f"it, a pixel skipping the trees behind it once its transmittance is <= "
            f"{f2b_stop:.5f}; a server frame after every pass")


def scale_bench(cells: int, frames: int, W: int, H: int, tile: int, B: int, pool_gb: float,
                rocks: int, max_trees: int, max_loading: int, serve_group: int,
                f2b_stop=None):
    """Served paging at scale: a (cells+1)^2 fBm + ridged terrain built tile
    by tile on the device (1 m cells), rocks instanced icospheres, tiles
    voxelized on demand into a pool of pool_gb GB, two camera streams."""
    import numpy as np
    from voxvizthesis_amd import scenes
    from voxvizthesis_amd.terrain import TerrainPager

    slots = wide_slots(B, pool_gb)
    t0 = time.perf_counter()
    pg = TerrainPager(cells, tile, B, slots, seed=6, ridged_seed=7, ridged_amp=0.05,
                      rocks=rocks, rock_seed=8, rock_radius=(2.0, 40.0), max_trees=max_trees,
                      max_loading=max_loading, streams=2, max_width=W, max_height=H,
                      near=1.0, far=4.0 * cells, serve_group=serve_group, f2b_stop=f2b_stop,
                      pool_normals=False)
    t_setup = time.perf_counter() - t0
    log(f"[bench] scale: {len(pg.trees)} paged trees set up in {t_setup:.1f} s")
    paths = [scenes.flythrough(seed=s, frames=frames, extent=float(cells), alt=(0.08, 0.2),
                               height_fn=pg.height_fn) for s in (6, 16)]
    try:
        wall, total, loaded_at = run_paged(pg, paths, frames, W, H)
        out = {
            "workload": f"{cells + 1}^2 fBm + ridged terrain (seeds 6/7, {cells} m, 1 m voxels) "
                        f"in a {cells}^3 virtual grid, {rocks} instanced icosphere rocks, "
                        f"{tile}^3 tiles of {B}^3 bricks built and voxelized on demand, two "
                        f"{W}x{H} camera streams over {frames}-frame seeded flythroughs",
            "pool_slots": int(slots),
            "pool_gb": round(float(slots) * (B + 2) ** 3 * 4 / 1e9, 2),
            "pool_format": "wide (slot in word 1), RGBA8 colour only",
            "cast": cast_mode(f2b_stop),
            "max_uploads_per_frame": 65536,
            "scene_setup_s": round(t_setup, 2),
            "frames": frames,
        }
        out.update(paged_stats(pg, wall, total, loaded_at))
        return out
    finally:
        pg.close()


def batch_bench(n_jobs: int, tmax: float, grids, do_autotune: bool, reps: int = 5):
    """Batched voxelization: n_jobs independent meshes (half fBm heightfield
    tiles, half random soups, triangle counts log-uniform 2K..tmax, grids from
    `grids`) captured into hipGraphs.  The autotuner picks each job's brick
    edge (8/16/32) and persistent launch shape by timing eager runs, and each
    graph's voxelize launch structure by timing replays: every job's own
    launches, or one persistent launch from a device work queue over all
    jobs or over the small ones.  Jobs are admitted by free device memory: while they fit, they join the current
    graph; when setting one up runs out of memory, the resident jobs are
    captured into one graph, timed and freed, and the next graph starts with
    that job (every job keeps its working buffers while its graph is
    resident).  Times are summed over the graphs."""
    import numpy as np
    from voxvizthesis_amd import scenes
    from voxvizthesis_amd.batch import Batch, autotune, make_job

    t0 = time.perf_counter()
    jobs_sc = scenes.batch_scenes(seed=4, n_jobs=n_jobs, tmax=tmax, grids=grids)
    log(f"[bench] batch: {n_jobs} meshes generated in {time.perf_counter() - t0:.1f} s")
    choices, t_tune = [], 0.0
    eager_ms = graph_ms = device_ms = check_ms = 0.0
    nodes, graphs, done, vox_launches = 0, 0, 0, 0
    # (jobs in the persistent launch, its workgroups; 0 = 4 per CU)
    vox_modes = ((0, 0), (-1, 0), (4096, 0))
    mode_ms, chosen = {}, []

    def oom(e):
        return "out of memory" in str(e)

    pending = list(jobs_sc)
    while pending:
        jobs = []
        t0 = time.perf_counter()
        while pending:
            sc = pending[0]
            try:
                if do_autotune:
                    B, pb, _, v = autotune(sc, reps=1, keep=True)
                else:
                    B, pb = 8, 0
                    v = make_job(sc, B, pb)
            except RuntimeError as e:
                if not oom(e) or not jobs:
                    raise
                break
            jobs.append(v)
            choices.append((B, pb))
            pending.pop(0)
            done += 1
            if done % 16 == 0:
                log(f"[bench] batch: {done} of {n_jobs} jobs tuned and set up")
        t_tune += time.perf_counter() - t0
        # eager: every job's setup + generate back to back (host round trips)
        eager = []
        for _ in range(reps):
            t1 = time.perf_counter()
            for v in jobs:
                v.computeEdgesFaceNormalsAndBounds()
                v.generateVoxelsAndOctTree(None)
            eager.append(time.perf_counter() - t1)
        bt = Batch(jobs)
        # the voxelize launch structure is tuned per graph: a launch per job
        # (0), one persistent launch over all jobs (-1), or over the jobs of
        # at most n active level-0 tiles (n) with the others launching their
        # own; the fastest (device time) is captured again and timed below
        tried = {}
        for K in vox_modes:
            bt.set_voxelize(persistent=K[0], blocks=K[1])
            bt.capture()
            dv = []
            for _ in range(reps):
                if not bt.launch():
                    raise RuntimeError(bt.getErrorMessage())
                dv.append(bt.stats().ms_graph_last)
            tried[K] = float(np.median(dv))
            mode_ms[K] = mode_ms.get(K, 0.0) + tried[K]
        best = min(tried, key=tried.get)
        chosen.append(best)
        bt.set_voxelize(persistent=best[0], blocks=best[1])
        bt.capture()
        graph, check, dev = [], [], []
        for _ in range(reps):
            t1 = time.perf_counter()
            if not bt.launch():
                raise RuntimeError(bt.getErrorMessage())
            graph.append(time.perf_counter() - t1)
            check.append(bt.stats().ms_check_last)
            dev.append(bt.stats().ms_graph_last)
        st = bt.stats()
        eager_ms += float(np.median(eager)) * 1e3
        graph_ms += float(np.median(graph)) * 1e3
        device_ms += float(np.median(dev))
        check_ms += float(np.median(check))
        nodes += int(st.graph_nodes)
        vox_launches += int(st.voxelize_launches)
        graphs += 1
        bt.close()
        for v in jobs:
            v.close()
        if pending:
            log(f"[bench] batch: graph {graphs} of {len(jobs)} jobs done, "
                f"{len(pending)} jobs left")
    T = sum(sc.ntri for sc in jobs_sc)
    vox = sum(sc.N[0] ** 3 for sc in jobs_sc)
    g = graph_ms * 1e-3
    return {
        "workload": f"{n_jobs} meshes (fBm tiles + random soups), triangles log-uniform "
                    f"2K-{int(tmax)}, grids {list(grids)}",
        "jobs": n_jobs,
        "triangles": int(T),
        "voxels": int(vox),
        "autotuned": bool(do_autotune),
        "brick_edges": {str(k): sum(1 for c in choices if c[0] == k) for k in (8, 16, 32)},
        "autotune_s": round(t_tune, 2),
        "graphs": graphs,
        "eager_ms": round(eager_ms, 3),
        "graph_ms": round(graph_ms, 3),
        "graph_device_ms": round(device_ms, 3),
        # device time summed over the graphs for each voxelize launch
        # structure tried, (jobs, workgroups): jobs 0 a launch per job, -1 one
        # persistent launch over all jobs, n over the jobs of at most n
        # level-0 tiles; workgroups of the persistent launch, 0 = 4 per CU;
        # and the structure chosen per graph
        "graph_device_ms_by_voxelize_launches": {str(k): round(v, 3)
                                                  for k, v in mode_ms.items()},
        "voxelize_mode_per_graph": chosen,
        "voxelize_launches": vox_launches,
        "checksum_ms": round(check_ms, 3),
        "graph_nodes": nodes,
        "meshes_per_s": round(n_jobs / g, 1),
        "gvoxels_per_s": round(vox / g / 1e9, 3),
        "mtriangles_per_s": round(T / g / 1e6, 2),
    }


def cpu_baseline(scene, win: int = 256):
    """The voxelization workload itself on the host cores: the oracle (the
    scalar C++ restatement of the reference's voxelizer) over the win^3
    windows that tile the workload's grid, at its own voxel size, each given
    the triangles whose bounds meet its voxel space (empty windows skipped),
    one window per thread at a time on every host core the process may use,
    at most 16 (the host share of one GPU).  Gvoxels/s = the whole grid over
    the wall time; the single-core equivalent divides by the summed window
    times instead."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import oracle
    from voxvizthesis_amd import scenes

    oracle.build()
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    N = scene.N[0]
    n = N // win
    d = np.float32(scene.delta[0])
    extra = (win // scene.B[0]) >> 1
    bounds = [(m.verts.min(axis=1), m.verts.max(axis=1)) for m in scene.meshes]
    jobs = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                p = (np.asarray(scene.p, np.float32) +
                     np.array([i, j, k], np.float32) * np.float32(win) * d).astype(np.float32)
                lo = p - np.float32(extra + 1) * d
                hi = p + np.float32(win + 3 * extra + 1) * d
                meshes = []
                for m, (bmin, bmax) in zip(scene.meshes, bounds):
                    sel = np.all((bmax >= lo) & (bmin <= hi), axis=1)
                    if sel.any():
                        meshes.append(scenes.Mesh(
                            verts=m.verts[sel],
                            vnormals=None if m.vnormals is None else m.vnormals[sel],
                            uvs=None if m.uvs is None else m.uvs[sel], image=m.image,
                            is_terrain=m.is_terrain, addr=m.addr))
                if meshes:
                    jobs.append(scenes.Scene(meshes=meshes, N=(win,) * 3, B=scene.B,
                                             p=tuple(float(x) for x in p), delta=scene.delta))

    def one(sc):
        t0 = time.perf_counter()
        r = oracle.run(sc, keep_mips=False)
        if r.status == 0:
            raise RuntimeError("oracle window failed")
        return time.perf_counter() - t0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        times = list(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    return {
        "value": round(N ** 3 / wall / 1e9, 6),
        "unit": "Gvoxels/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"the whole {N}^3 workload ({scene.name}) as {len(jobs)} of its {n ** 3} "
                   f"windows of {win}^3 voxels (the others hold no triangle), each voxelized "
                   f"by the oracle (scalar C++ port of the reference logic, {win // scene.B[0]}"
                   f"-brick windows) on one of {cores} threads; {wall:.1f} s wall, "
                   f"{sum(times):.1f} s summed"),
        "seconds": round(wall, 3),
        "window_seconds_sum": round(sum(times), 3),
        "single_core_equivalent": round(N ** 3 / sum(times) / 1e9, 6),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied once by a plain copy kernel (profiled runs)")
    ap.add_argument("--cpu-window", type=int, default=256,
                    help="window edge of the host-core baseline over the workload")
    ap.add_argument("--serve-frames", type=int, default=300)
    ap.add_argument("--serve-width", type=int, default=1920)
    ap.add_argument("--serve-height", type=int, default=1080)
    ap.add_argument("--serve2-width", type=int, default=2560)
    ap.add_argument("--serve2-height", type=int, default=1440)
    ap.add_argument("--serve-max-uploads", type=int, default=65536)
    ap.add_argument("--serve-evict-slots", type=int, default=16,
                    help="slots per axis of the serving leg whose pool is smaller than the "
                         "octree's bricks")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--paged-cells", type=int, default=8192)
    ap.add_argument("--paged-pool-gb", type=float, default=16.0)
    ap.add_argument("--no-paged", action="store_true")
    ap.add_argument("--scale-cells", type=int, default=32768)
    ap.add_argument("--scale-pool-gb", type=float, default=128.0)
    ap.add_argument("--scale-rocks", type=int, default=100000)
    ap.add_argument("--scale-frames", type=int, default=300)
    ap.add_argument("--no-scale", action="store_true")
    ap.add_argument("--scale-f2b", action="store_true",
                    help="cast the scale leg's trees front to back in passes")
    ap.add_argument("--paged-trees", type=int, default=96,
                    help="octrees the paged leg keeps loaded (tree slots of the server)")
    ap.add_argument("--f2b-stop", type=float, default=1.0 / 255.0,
                    help="paged_front_to_back leg: transmittance at which a pixel stops")
    ap.add_argument("--no-paged-f2b", action="store_true",
                    help="skip the C4 leg cast front to back in passes")
    ap.add_argument("--no-paged-per-tree", action="store_true",
                    help="skip the C4 leg that casts every tree alone")
    ap.add_argument("--paged-group", type=int, default=2,
                    help="trees per server frame within a paged frame (0: one per frame)")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-tmax", type=float, default=2e6)
    ap.add_argument("--batch-grids", default="64,128,256,512,1024")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--chunk", type=int, default=256,
                    help="also time the workload in chunks of this edge (the reference's "
                         "default voxChunkDim); 0 skips it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the voxelizer "
              f"shards by tile across GPUs by design but this build measures one GPU.",
              file=sys.stderr)
        sys.exit(2)

    from voxvizthesis_amd import scenes
    from voxvizthesis_amd._lib import VOX_OPT_TIMING
    from voxvizthesis_amd.voxelizer import Voxelizer

    t0 = time.perf_counter()
    scene = scenes.scene_full_octree()
    log(f"[bench] scene {scene.name}: {scene.ntri} triangles, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    B = scene.B[0]
    vz = Voxelizer(scene.N, scene.B, None)
    ok = vz.initCuda() and vz.allocateVoxelMemory() and vz.setVoxelizationParams(scene.p,
                                                                              scene.delta)
    for m in scene.meshes:
        ok = ok and vz.addMesh(m)
    t_up = time.perf_counter()
    ok = ok and vz.allocateTriangleMemory()           # H2D once: inputs resident in HBM
    t_up = time.perf_counter() - t_up
    if not ok:
        raise RuntimeError(vz.getErrorMessage())

    def step():
        if not vz.computeEdgesFaceNormalsAndBounds():
            raise RuntimeError(vz.getErrorMessage())
        if vz.generateVoxelsAndOctTree(None) != 1:
            raise RuntimeError(vz.getErrorMessage())

    if args.copy_calibration:
        # one plain copy of known size, the reference point of the traffic
        # counters of a profiled run (not part of any timed figure)
        from voxvizthesis_amd import _lib
        L = _lib.load()
        nb = args.copy_calibration << 20
        src, dst = L.vox_device_alloc(nb), L.vox_device_alloc(nb)
        if not (src and dst and L.vox_device_copy(dst, src, nb)):
            raise RuntimeError("calibration copy failed")
        # the same bytes in 256-byte blocks visited in a scattered order: a
        # second reference point, closer to the gathering kernels
        if not L.vox_device_copy_scattered(dst, src, nb):
            raise RuntimeError("scattered calibration copy failed")
        L.vox_device_free(src)
        L.vox_device_free(dst)
        log(f"[bench] calibration copy of {nb} bytes done")
    for i in range(args.warmup):
        step()
    log(f"[bench] warmup {args.warmup} done")
    # generate() ends with a stream synchronize, so the timed region is
    # bracketed by device synchronisation on both sides.
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    st = vz.stats()
    res = vz.result(with_bricks=False)
    bricks = [t["nb"] for t in res["tree"]]
    log(f"[bench] {args.steps} steps: {ms:.3f} ms/step, bricks {bricks}")

    # one more step with per-kernel HIP events for the roofline
    vz._L.vox_set_options(vz._h, VOX_OPT_TIMING)
    step()
    ts = vz.stats()
    kern = {}
    for i in range(ts.nkernels):
        name = ts.kernel_name[i].value.decode()
        base = name.split("_g")[0] if name.startswith("voxelize_g") else name
        ms0, n0 = kern.get(base, (0.0, 0))
        kern[base] = (ms0 + ts.kernel_ms[i], n0 + ts.kernel_launches[i])
    kern_detail = {ts.kernel_name[i].value.decode(): round(ts.kernel_ms[i], 4)
                   for i in range(ts.nkernels)}
    vz._L.vox_set_options(vz._h, 0)
    N = scene.N[0]
    total_bytes = algorithmic_bytes(scene, st, bricks, B)
    # Measured HBM bytes per launch of a kernel (rocprofv3 FETCH_SIZE +
    # WRITE_SIZE passes over this workload, each scaled by the reading of a
    # plain copy of known size), written by
    # `python -m voxvizthesis_amd.profiling traffic ... --json profiles/traffic.json`.
    tj = {}
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f).get("kernels", {})

    def roofline_of(name):
        kms, launches = kern[name]
        pl = kms / launches
        # the stats of the timing step: it also counts the mip sources
        kb = kernel_bytes(name, scene, ts, bricks, B, launches)
        ach = kb / (pl * 1e-3) / 1e9
        # the kernel that does the work of every depth in one launch carries "_all"
        tr = tj.get("k_" + name) or tj.get("k_" + name + "_all")
        out = {
            "kernel": name,
            "bound": "hbm",
            "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": int(tr["bytes_per_launch"]) if tr else None,
            "per_launch_ms": round(pl, 4),
            "algorithmic_bytes_per_launch": int(kb),
        }
        if name == "voxelize":
            # fraction above is of HBM bandwidth; what bounds this kernel is
            # VALU issue (SQ counters, profiles/full_octree_traffic.md)
            out["limiter"] = "VALU issue: ~75 % of the SIMDs' VALU issue rate (SQ counters)"
        return out

    # The roofline names k_voxelize (one launch per triangle group), the
    # voxelization kernel and the largest kernel time per step in the rocprof
    # trace; the brick gather (within a few percent of it) and the mip filter
    # follow in "rooflines".
    out = {
        "metric": METRIC,
        "value": round(N ** 3 / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm terrain + icosphere rocks)",
        "config": {
            "workload": "full octree: 1025^2 fBm terrain + 2,000 rocks, 1024^3 grid, 16^3 "
                        "bricks, depths 0-6",
            "triangles": int(st.triangles),
            "grid": N,
            "brick": B,
            "levels": len(bricks),
        },
        "bricks_per_s": round(sum(bricks) / (ms * 1e-3), 1),
        "bricks": bricks,
        "pipeline_hbm_gbs": round(total_bytes / (ms * 1e-3) / 1e9, 1),
        "pipeline_hbm_frac": round(total_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "pipeline_algorithmic_bytes": int(total_bytes),
        "pcie_upload_s": round(t_up, 4),
        "kernels_ms": kern_detail,
        "roofline": roofline_of("voxelize"),
        "rooflines": [roofline_of(k) for k in ("brick_gather", "mip", "setup", "scan", "scatter")
                      if k in kern],
    }
    # an attainable-bandwidth reference beside the 8 TB/s peak: a plain
    # 16-byte-per-lane device copy of 4 GiB (read + write), median of 10;
    # left out of counter runs, whose one calibration copy must be the only
    # launch of the copy kernel
    if not args.copy_calibration:
        out["hbm_copy"] = copy_reference(4 << 30, 10)
        log(f"[bench] plain copy: {out['hbm_copy']['gbs']} GB/s")
    if args.chunk > 0:
        log(f"[bench] the same workload in {args.chunk}^3 chunks ...")
        out["chunked"] = chunked_bench(scene, args.chunk, args.steps, args.warmup)
        log(f"[bench] chunked: {out['chunked']['ms_per_step']} ms/step")
    if not args.no_serve:
        log(f"[bench] serving {args.serve_frames} frames at {args.serve_width}x"
            f"{args.serve_height} ...")
        sv = serving_bench(vz, scene, args.serve_frames, args.serve_width, args.serve_height,
                           (56, 56, 56), args.serve_max_uploads)
        out["serving"] = sv
        log(f"[bench] serving: {sv['frame_ms']} ms/frame, p99 {sv['latency_p99_ms']} ms")
        log(f"[bench] serving two {args.serve2_width}x{args.serve2_height} streams ...")
        sv2 = serving_bench(vz, scene, args.serve_frames, args.serve2_width, args.serve2_height,
                            (56, 56, 56), args.serve_max_uploads, streams=2)
        out["serving_two_streams"] = sv2
        log(f"[bench] two streams: {sv2['frame_ms']} ms/frame, p99 {sv2['latency_p99_ms']} ms")
        # a pool smaller than the octree's bricks, so the LRU evicts
        e = args.serve_evict_slots
        sv3 = serving_bench(vz, scene, args.serve_frames, args.serve_width, args.serve_height,
                            (e, e, e), args.serve_max_uploads)
        out["serving_evicting"] = sv3
        log(f"[bench] evicting pool: {sv3['frame_ms']} ms/frame, {sv3['evictions']} evictions, "
            f"p99 {sv3['latency_p99_ms']} ms")
    if not args.no_paged and not args.no_paged_per_tree:
        log(f"[bench] paged serving over a {args.paged_cells}^3 virtual grid ...")
        pb = paged_bench(args.paged_cells, args.serve_frames, args.serve_width,
                         args.serve_height, 1024, 16, args.paged_pool_gb, args.paged_trees, 4,
                         args.paged_group)
        out["paged"] = pb
        log(f"[bench] paged: {pb['frame_ms']} ms/frame, {pb['tiles_voxelized']} tiles, "
            f"p99 {pb['latency_p99_ms']} ms")
    if not args.no_paged and not args.no_paged_f2b:
        log("[bench] the same, cast front to back in passes ...")
        pf = paged_bench(args.paged_cells, args.serve_frames, args.serve_width,
                         args.serve_height, 1024, 16, args.paged_pool_gb, args.paged_trees, 4,
                         args.paged_group, args.f2b_stop)
        out["paged_front_to_back"] = pf
        log(f"[bench] front to back: {pf['frame_ms']} ms/frame, p99 {pf['latency_p99_ms']} ms")
    if not args.no_scale:
        log(f"[bench] paged serving at scale over a {args.scale_cells}^3 virtual grid ...")
        sb = scale_bench(args.scale_cells, args.scale_frames, args.serve2_width,
                         args.serve2_height, 1024, 16, args.scale_pool_gb, args.scale_rocks,
                         args.paged_trees, 4, args.paged_group,
                         args.f2b_stop if args.scale_f2b else None)
        out["paged_scale"] = sb
        log(f"[bench] scale: {sb['frame_ms']} ms/frame, {sb['tiles_voxelized']} tiles, "
            f"p99 {sb['latency_p99_ms']} ms")
    if not args.no_batch:
        log(f"[bench] batch of {args.batch_jobs} meshes ...")
        bb = batch_bench(args.batch_jobs, args.batch_tmax,
                         tuple(int(x) for x in args.batch_grids.split(",")),
                         not args.no_autotune)
        out["batched"] = bb
        log(f"[bench] batch: graph {bb['graph_ms']} ms vs eager {bb['eager_ms']} ms")
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline: the workload in {args.cpu_window}^3 windows ...")
        cb = cpu_baseline(scene, args.cpu_window)
        out["cpu_baseline"] = cb
        out["gpu_vs_cpu"] = round(out["value"] / cb["value"], 1)
        out["gpu_vs_cpu_single_core"] = round(out["value"] / cb["single_core_equivalent"], 1)
        log(f"[bench] cpu: {cb['value']} Gvoxels/s on {cb['cores']} cores "
            f"({cb['single_core_equivalent']} single-core equivalent)")
    vz.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
