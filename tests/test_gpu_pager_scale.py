"""Paged serving at a larger scale (GPU): a 2049^2 fBm + ridged terrain with
2,000 rocks in 512^3 tiles of 16^3 bricks (the paged workloads' brick edge,
tiles an eighth of theirs by volume), voxelized on demand over a 16-frame
flythrough into a colour-only wide pool too small for every brick, served
front to back in passes of two trees.  Every loaded tile equals the same
triangles voxelized alone, bit for bit, and a sample of the pool's resident
slots each hold the brick of the node they serve."""
import numpy as np
import pytest

from oracle.server import NONE
from voxvizthesis_amd import scenes
from voxvizthesis_amd.terrain import TerrainPager
from tests.test_gpu_terrain import alone, snapshot

pytestmark = pytest.mark.gpu


def test_paged_tiles_at_512_cubed_equal_alone_and_pool_is_consistent():
    W, H = 320, 180
    slots = 4000    # 93 MB of 18^3 colour bricks, fewer than the frames request
    pg = TerrainPager(2048, 512, 16, slots, seed=6, ridged_seed=7, ridged_amp=0.05, rocks=2000,
                      rock_radius=(2.0, 20.0), max_trees=10, max_loading=3, streams=1,
                      max_width=W, max_height=H, near=1.0, far=8000.0, serve_group=2,
                      f2b_stop=1.0 / 255.0, pool_normals=False)
    try:
        print(f"[scale] {len(pg.trees)} paged trees", flush=True)
        path = scenes.flythrough(seed=5, frames=16, extent=2048.0, alt=(0.05, 0.15),
                                 height_fn=pg.height_fn)
        for f in range(16):
            if f in (3, 7, 11):
                pg.drain()
            pg.frame([dict(position=path[f][0], look=path[f][1], up=path[f][2], fov_y_deg=60.0,
                           width=W, height=H)])
        pg.drain()
        pg.frame([dict(position=path[-1][0], look=path[-1][1], up=path[-1][2], fov_y_deg=60.0,
                       width=W, height=H)])
        loaded = [t for t in pg.trees if t.state == "loaded"]
        st = pg.server.stats()
        print(f"[scale] loaded {len(loaded)} trees, {st.uploads} uploads, {st.evictions} "
              f"evictions, load ms {np.round(pg.load_ms, 1)}", flush=True)
        assert len(loaded) >= 3
        assert st.evictions > 0    # the LRU had to make room
        for tr in loaded[:3]:
            st_, ref = alone(pg, tr)
            assert st_ == 1
            got = snapshot(tr.vz)
            assert len(got) == len(ref) and all((a == b).all() for a, b in zip(got, ref)), \
                f"tile {tr.x},{tr.y},{tr.z}"
        # a sample of resident slots hold their nodes' bricks
        node_slot, slot_node = pg.server.state()
        resident = np.nonzero(np.asarray(slot_node) != NONE)[0]
        assert resident.size > 0
        rng = np.random.Generator(np.random.PCG64(1))
        sample = rng.choice(resident, size=min(300, resident.size), replace=False)
        view = pg.pool.view()
        e3 = 18 ** 3
        by_slot = {pg.server_slot(t): t for t in loaded}
        info_of = {}
        for s in sample:
            k, x = divmod(int(slot_node[s]), pg.cap)
            tr = by_slot[k]
            if k not in info_of:
                res = tr.vz.result()
                info = {}
                for d in range(res["levels"]):
                    tx = tr.octree.node_texel(d)
                    for nn in np.nonzero(tx != NONE)[0]:
                        info[int(tx[nn])] = (d, int(nn))
                info_of[k] = (res, info)
            res, info = info_of[k]
            d, nn = info[x]
            b = int(res["tree"][d]["bidx"][nn])
            got = np.empty(e3, np.uint32)
            assert pg._L.vox_pool_copy_to_host(got.ctypes.data, view.colors + 4 * e3 * int(s),
                                               got.nbytes)
            assert (got == res["tree"][d]["bcol"][b]).all(), f"slot {s}"
        print(f"[scale] {sample.size} sampled slots equal their bricks", flush=True)
    finally:
        pg.close()
