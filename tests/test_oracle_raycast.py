"""Pins the ray-cast oracle (oracle/ref_raycast.cpp) on the CPU with cases
whose outcome follows from the shader's rules directly: rays that miss the
box record nothing, a transparent constant root is recorded once by the
pixels whose 2x2 push window opens first, an opaque constant root composites
to its own colour, a uniform brick accumulates alpha 1-(1-a)^n with colour
proportional to alpha, and the level of detail descends only when the
projected voxel exceeds the pixel size."""
import numpy as np

import oracle

CB = 0x40000000
RB = 0x80000000
ROOT_ID = (1 << 20) | (1 << 10) | 1


def pk(x, y, z):
    return (x << 20) | (y << 10) | z


def cam(pos, look, up=(0, 1, 0), w=32, h=24, fov=40.0, frame=0):
    return dict(position=pos, look=look, up=up, fov_y_deg=fov, width=w, height=h, frame=frame)


# a short ray step keeps the jittered start within 1e-3 of the box face
PRM = dict(vol_min=(0, 0, 0), vol_max=(1, 1, 1), jitter_seed=7, ray_step=1e-3)


def ray_hits_box(c, margin):
    """Analytic (double) ray / unit-box test per pixel with a margin."""
    W, H = c["width"], c["height"]
    look, up = np.array(c["look"], float), np.array(c["up"], float)
    left = np.cross(up, look)
    ty = np.tan(np.radians(c["fov_y_deg"]) / 2)
    tx = ty * W / H
    pos = np.array(c["position"], float)
    hit = np.zeros((H, W), int)  # 1 hit, 0 miss, -1 near the silhouette
    for j in range(H):
        for i in range(W):
            sx = 2 * (i + 0.5) / W - 1
            sy = 2 * (j + 0.5) / H - 1
            d = look - left * tx * sx + up * ty * sy
            t0 = (0 - margin - pos) / d
            t1 = (1 + margin - pos) / d
            big = np.maximum(np.minimum(t0, t1).max(), 0) <= np.maximum(t0, t1).min()
            t0 = (0 + margin - pos) / d
            t1 = (1 - margin - pos) / d
            small = np.maximum(np.minimum(t0, t1).max(), 0) <= np.maximum(t0, t1).min()
            hit[j, i] = 1 if small else (0 if not big else -1)
    return hit


def one_node_tree(w0, w1):
    nodes = np.zeros((2, 2, 2, 2), np.uint32)
    nodes[0, 0, 0] = (w0, w1)
    return nodes


def test_transparent_constant_root_recorded_by_first_window():
    nodes = one_node_tree(CB, 0)
    pool = np.zeros((4, 4, 4), np.uint32)
    c = cam((0.5, 0.5, 3.0), (0, 0, -1))
    usage, rgba = oracle.raycast(c, PRM, 1, (2, 2, 2), nodes, pool)
    hit = ray_hits_box(c, 3e-3)
    assert (hit == 1).sum() > 50 and (hit == 0).sum() > 50
    H, W = hit.shape
    for j in range(H):
        for i in range(W):
            u = usage[j, i]
            if hit[j, i] == 0:
                assert not u.any()
            elif hit[j, i] == 1:
                first = (i % 2 == 0 and j % 2 == 0)
                assert (u[0] == ROOT_ID) == first and not u[1:].any()
    assert not rgba.any()


def test_opaque_constant_root_composites_its_colour():
    col = (200, 100, 50, 255)
    w1 = col[0] | (col[1] << 8) | (col[2] << 16) | (col[3] << 24)
    nodes = one_node_tree(CB, w1)
    pool = np.zeros((4, 4, 4), np.uint32)
    c = cam((0.5, 0.5, 3.0), (0, 0, -1))
    usage, rgba = oracle.raycast(c, PRM, 1, (2, 2, 2), nodes, pool)
    hit = ray_hits_box(c, 3e-3)
    exp = 0
    for k, v in enumerate(col):
        f = np.float32(np.float32(v) / np.float32(255.0)) * np.float32(255.0)
        exp |= int(np.float32(f)) << (8 * k)
    assert (rgba[hit == 1] == exp).all()
    assert (rgba[hit == 0] == 0).all()


def test_uniform_brick_accumulates_alpha():
    B = 2
    e = B + 2
    a8 = 20
    col = (10, 5, 2, a8)      # not premultiplied: colour / alpha stays below 1
    pool = np.full((e, e, e), col[0] | (col[1] << 8) | (col[2] << 16) | (col[3] << 24),
                   np.uint32)
    nodes = one_node_tree(RB, pk(1, 1, 1))
    c = cam((0.5, 0.5, 3.0), (0, 0, -1), w=8, h=8, fov=2.0)
    prm = dict(PRM, ray_step=0.01)
    usage, rgba = oracle.raycast(c, prm, 1, (B, B, B), nodes, pool)
    a = a8 / 255.0
    for j in range(3, 5):
        for i in range(3, 5):
            v = int(rgba[j, i])
            r, alpha = (v & 255) / 255.0, (v >> 24) / 255.0
            # colour stays proportional to alpha: r = (10/255)/a * alpha
            assert abs(r - (10 / 255.0) / a * alpha) < 3 / 255.0
            # about 100 samples over the unit depth, but compositing stops at
            # the first alpha above 0.995: 1 - (1-a)^k for the smallest such k
            k = int(np.ceil(np.log(0.005) / np.log(1 - a)))
            assert k < 98
            assert abs(alpha - (1 - (1 - a) ** k)) <= 1.5 / 255.0


def two_level_tree():
    """Root NON-CONST with a resident brick, children CONST and opaque, each
    with its own colour."""
    nodes = np.zeros((2, 2, 4, 2), np.uint32)     # 16 nodes: dims (4, 2, 2)
    nodes[0, 0, 0] = (RB | pk(2, 0, 0), pk(1, 1, 1))
    for k in range(8):
        x, y, z = 2 + (k & 1), (k >> 1) & 1, k >> 2
        nodes[z, y, x] = (CB, (10 * (k + 1)) | (0xFF << 24))
    return nodes


def test_level_of_detail_descends_only_when_close():
    nodes = two_level_tree()
    pool = np.zeros((4, 4, 4), np.uint32)
    pool[:] = 0x80808080
    near = cam((0.5, 0.5, 1.6), (0, 0, -1), w=16, h=16, fov=60.0)
    far = cam((0.5, 0.5, 5000.0), (0, 0, -1), w=16, h=16, fov=0.02)
    un, _ = oracle.raycast(near, PRM, 2, (2, 2, 2), nodes, pool)
    uf, _ = oracle.raycast(far, PRM, 2, (2, 2, 2), nodes, pool)
    ids_near = set(un[un != 0].tolist())
    ids_far = set(uf[uf != 0].tolist())
    child_ids = {((2 + (k & 1) + 1) << 20) | ((((k >> 1) & 1) + 1) << 10) | ((k >> 2) + 1)
                 for k in range(8)}
    assert ids_far == {ROOT_ID}
    assert ids_near and ids_near <= child_ids


def test_wide_pool_slot_equals_a_one_slot_reference_pool():
    """A wide pool (word 1 of a resident node is its slot, each brick sampled
    in its own frame) reads slot s exactly as the reference's pool texture
    of a single slot holding that brick reads it: same usage, same colour.
    A word 1 that names no slot reads slot 0."""
    rng = np.random.default_rng(3)
    e = 4   # B = 2
    bricks = rng.integers(0, 1 << 32, size=(3, e, e, e), dtype=np.uint64).astype(np.uint32)
    bricks |= np.uint32(0x60000000)   # alpha well above zero
    wide_pool = bricks.reshape(3 * e, e, e)            # slots stacked along z
    c = cam((0.45, 0.55, 2.5), (0, 0, -1), w=24, h=16)
    for s in range(3):
        ref_u, ref_c = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, pk(1, 1, 1)), bricks[s])
        w_u, w_c = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, s), wide_pool, wide=True)
        assert (ref_u == w_u).all() and (ref_c == w_c).all()
        assert ref_c.any()
    # an out-of-range word 1 reads slot 0
    ref_u, ref_c = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, pk(1, 1, 1)), bricks[0])
    w_u, w_c = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, 7), wide_pool, wide=True)
    assert (ref_c == w_c).all()
    # and the slots differ from each other
    _, c1 = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, 1), wide_pool, wide=True)
    _, c2 = oracle.raycast(c, PRM, 1, (2, 2, 2), one_node_tree(RB, 2), wide_pool, wide=True)
    assert (c1 != c2).any()
