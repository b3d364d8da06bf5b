"""HIP ray cast against the ray-cast oracle, bit-exact per pixel (GPU): the
12 node ids every pixel records and the composited RGBA8 colour, for several
camera poses over voxelized scenes; the device request list holds exactly
the distinct recorded ids."""
import numpy as np
import pytest

import oracle
from voxvizthesis_amd import scenes
from voxvizthesis_amd._lib import VOX_OPT_MIP_CHAIN
from voxvizthesis_amd.gvpool import BrickPool, OctTree
from voxvizthesis_amd.gvserve import RayCaster, camera, raycast_params
from voxvizthesis_amd.voxelizer import voxelize_scene

pytestmark = pytest.mark.gpu


def look_at(pos, target, up=(0.0, 0.0, 1.0)):
    pos, target, up = (np.asarray(v, np.float64) for v in (pos, target, up))
    look = target - pos
    look /= np.linalg.norm(look)
    up = up - up.dot(look) * look
    up /= np.linalg.norm(up)
    return look.astype(np.float32), up.astype(np.float32)


POSES = [
    ((0.5, -0.6, 1.3), (0.5, 0.5, 0.5), 45.0, 0),
    ((1.4, 1.2, 0.9), (0.4, 0.5, 0.45), 35.0, 1),
    ((0.3, 0.4, 0.62), (0.7, 0.6, 0.5), 70.0, 2),      # inside the box, near the surface
]


def _frames(scene, W, H, options=VOX_OPT_MIP_CHAIN, slots=None, seed=3):
    print(f"[raycast] {scene.name} {W}x{H}", flush=True)
    vz = voxelize_scene(scene, options=options)
    bp = BrickPool(scene.B, slots)
    t = OctTree(bp, vz)
    rc = RayCaster(t, W, H)
    try:
        nodes = t.nodes()
        pool = bp.textures()[0]
        L = vz.raw_result().levels
        lo = np.asarray(scene.p, np.float32)
        hi = lo + np.asarray(scene.N, np.float32) * np.asarray(scene.delta, np.float32)
        for pos, target, fov, frame in POSES:
            look, up = look_at(pos, target)
            cam = dict(position=pos, look=look, up=up, fov_y_deg=fov, width=W, height=H,
                       frame=frame)
            prm = dict(vol_min=lo, vol_max=hi, jitter_seed=seed)
            print(f"[raycast] pose {pos} frame {frame}", flush=True)
            v = rc.frame(camera(**cam), raycast_params(**prm))
            gu, gc = rc.usage(), rc.rgba()
            ru, rcol = oracle.raycast(cam, prm, L, scene.B, nodes, pool)
            bad = np.argwhere((gu != ru).any(axis=-1))
            assert bad.size == 0, f"{len(bad)} pixels' usage differ, first {bad[:4].tolist()}"
            badc = np.argwhere(gc != rcol)
            assert badc.size == 0, f"{len(badc)} pixels' colour differ, first {badc[:4].tolist()}"
            ids = set(ru[ru != 0].tolist())
            req = rc.requests()
            assert v.nrequests == v.nrequests_total == len(ids)
            assert len(set(req.tolist())) == len(req) and set(req.tolist()) == ids
            yield ru, rcol
    finally:
        rc.close()
        t.close()
        bp.close()
        vz.close()


def test_raycast_small_terrain():
    n = 0
    for ru, rcol in _frames(scenes.scene_small_terrain(), 96, 64):
        n += int((ru != 0).sum())
    assert n > 0


def test_raycast_mixed_groups_reference_mips():
    for _ in _frames(scenes.scene_mixed(), 64, 48, options=0, slots=(4, 4, 4)):
        pass
