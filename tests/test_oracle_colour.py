"""Pins the oracle's colour path away from degenerate settings (CPU only).

ComputeColorsAndNormals / ComputeColorsAsNormals (VoxelizerKernels.cu
602-655, 835-849, 505-514, 689-1083) are checked against closed forms
evaluated in float64, on inputs where a plausible slip changes the answer
by several quantisation steps:

* affine vertex normals and UVs over a linear-ramp texture, voxel centres
  off the triangle's plane: the normal and the texture coordinate are the
  affine functions at the clamped barycentric coordinates of the voxel
  centre projected along the dominant axis; the colour is the bilinear
  filter of the ramp with the texel-centre offset of -0.5 and weights held
  in 8 fractional bits (a swapped nB / nC, swapped u / v or a missing -0.5
  each move a channel by 4 steps or more);
* seven triangles through one voxel, each with its own texel: the slot rule
  (three lowest indices plus the highest, VK:505-514 linearised) against the
  mean of the first four, of all seven, or of other subsets;
* an untextured group with vertex normals and a negative dominant normal
  component (the signed axis choice of VK:1023-1030, not the absolute one of
  the textured kernel), and seven untextured triangles on one voxel (the sum
  over four slots divided by the raw hit count, alpha min(4, 1) undivided,
  VK:997, 1066-1068).

Hit sets per voxel come from the float64 separating-axis test of
test_oracle.py, so voxels in its tie band are skipped.
"""
import itertools

import numpy as np

import oracle
from tests.test_oracle import occ_grid, sat_margin, voxel_bbox_f32
from voxvizthesis_amd import scenes

N, B = 32, 8


def _grid_params(sc):
    dp = np.float32(sc.delta[0])
    extra = (N // B) >> 1
    offp = np.float32(sc.p[0]) - np.float32(extra) * dp
    return offp, dp


def _centre(offp, dp, x, y, z):
    """Voxel centre as the reference computes it (origin + i*delta + delta/2),
    in float64 from the float32 inputs."""
    o, d = float(offp), float(dp)
    return np.array([o + x * d + 0.5 * d, o + y * d + 0.5 * d, o + z * d + 0.5 * d])


def _bary_clamped(P, A, Bv, Cv):
    """ComputeVoxelBarycentricCoords in float64: (u, v) with P ~ A + u (C - A)
    + v (B - A), each clamped to [0, 1] on its own."""
    v0, v1, v2 = Cv - A, Bv - A, P - A
    d00, d01, d02 = v0 @ v0, v0 @ v1, v0 @ v2
    d11, d12 = v1 @ v1, v1 @ v2
    den = d00 * d11 - d01 * d01
    u = (d11 * d02 - d01 * d12) / den
    v = (d00 * d12 - d01 * d02) / den
    return min(max(u, 0.0), 1.0), min(max(v, 0.0), 1.0)


_PROJ = {0: (2, 1), 1: (0, 2), 2: (0, 1)}   # dominant axis -> projected coordinates


def _hit_sets(verts, sc):
    """Per voxel (x, y, z) the sorted triangles that overlap it in float64, and
    the voxels in the tie band."""
    offp, dp = _grid_params(sc)
    S = [n + 4 * ((N // B) >> 1) for n in sc.N]
    hits, ties = {}, set()
    for t in range(verts.shape[0]):
        bmin, bmax = voxel_bbox_f32(verts[t], offp, dp, S)
        tri = verts[t].astype(np.float64)
        for z, y, x in itertools.product(range(bmin[2], bmax[2] + 1), range(bmin[1], bmax[1] + 1),
                                         range(bmin[0], bmax[0] + 1)):
            lo = float(offp) + np.array([x, y, z], np.float64) * float(dp)
            m = sat_margin(tri, lo, lo + float(dp))
            if abs(m) < 1e-5 * float(dp):
                ties.add((x, y, z))
            elif m <= 0:
                hits.setdefault((x, y, z), []).append(t)
    return hits, ties


def _slots(h):
    """Slot triangles of a voxel hit by the sorted list h (three lowest, then
    the highest once four or more hit)."""
    return h[:3] + ([h[-1]] if len(h) >= 4 else [])


def _chan(c, k):
    return (int(c) >> (8 * k)) & 255


def _within_one_step(got, exact):
    # truncation of the float32 value: floor(exact) or one below it
    return exact - 1.0 - 1e-3 <= got <= exact + 1e-3


# ---------------------------------------------------------------------------
def _face_normal(v):
    n = np.cross(v[1] - v[0], v[2] - v[0])
    return n / np.linalg.norm(n)


def test_affine_normals_and_uvs_over_a_ramp():
    W = 32
    img = np.zeros((1, W, 4), np.uint8)
    img[0, :, 0] = 8 * np.arange(W)           # R ramp: 8 per texel
    img[0, :, 1] = 255 - 8 * np.arange(W)     # G falling ramp
    img[0, :, 2] = 100
    img[0, :, 3] = 255
    verts = np.array([[[0.21, 0.22, 0.40], [0.71, 0.25, 0.47], [0.33, 0.78, 0.45]]], np.float32)
    vn = np.array([[[0.10, 0.20, 0.97], [-0.30, 0.10, 0.95], [0.20, -0.40, 0.89]]], np.float32)
    uvs = np.array([[[0.10, 0.5], [0.85, 0.5], [0.40, 0.5]]], np.float32)
    sc = scenes.Scene(meshes=[scenes.Mesh(verts=verts, vnormals=vn, uvs=uvs, image=img)],
                      N=(N,) * 3, B=(B,) * 3)
    ref = oracle.run(sc)
    col, nrm = ref.mips[0]
    occ = occ_grid(ref)
    offp, dp = _grid_params(sc)
    hits, ties = _hit_sets(verts, sc)
    v = verts[0].astype(np.float64)
    fn = _face_normal(v)
    axis = int(np.argmax(np.abs(fn)))
    assert axis == 2
    a, b = _PROJ[axis]
    nA, nB, nC = vn[0].astype(np.float64)
    tA, tB, tC = uvs[0].astype(np.float64)
    T = img[0, :, :3].astype(np.float64)
    checked = 0
    for (x, y, z), h in hits.items():
        if (x, y, z) in ties:
            continue
        assert occ[z, y, x]
        c = _centre(offp, dp, x, y, z)
        u, w = _bary_clamped(c[[a, b]], v[0][[a, b]], v[1][[a, b]], v[2][[a, b]])
        n = nA + u * (nC - nA) + w * (nB - nA)
        n /= np.linalg.norm(n)
        assert np.allclose(nrm[z, y, x], n, atol=2e-5), ((x, y, z), nrm[z, y, x], n)
        t = tA + u * (tC - tA) + w * (tB - tA)
        xb = t[0] * W - 0.5
        i = int(np.floor(xb))
        a8 = np.round((xb - i) * 256.0) / 256.0
        exact = (1.0 - a8) * T[i] + a8 * T[i + 1]
        got = col[z, y, x]
        for k in range(3):
            assert _within_one_step(_chan(got, k), exact[k]), ((x, y, z), k, _chan(got, k), exact[k])
        assert _chan(got, 3) == 255
        checked += 1
    assert checked > 100
    # every occupied voxel is one of the float64 hits or in the tie band
    zz, yy, xx = np.nonzero(occ)
    assert all((x, y, z) in hits or (x, y, z) in ties for x, y, z in zip(xx, yy, zz))


def _fan(centre, k, n_tris, radius, seed):
    """n_tris triangles in random planes through `centre`, each with the centre
    as its centroid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tris = []
    for _ in range(n_tris):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        u1 = np.cross(n, [0.3, 0.5, 0.8])
        u1 /= np.linalg.norm(u1)
        u2 = np.cross(n, u1)
        ang = rng.uniform(0, 2 * np.pi)
        pts = [centre + radius * (np.cos(ang + j * 2 * np.pi / 3) * u1 +
                                  np.sin(ang + j * 2 * np.pi / 3) * u2) for j in range(3)]
        tris.append(pts)
    return np.array(tris, np.float32)


def test_slot_rule_three_lowest_and_highest():
    sc0 = scenes.Scene(meshes=[], N=(N,) * 3, B=(B,) * 3)
    offp, dp = _grid_params(sc0)
    target = (16, 16, 16)
    c0 = _centre(offp, dp, *target)
    n_t = 7
    verts = _fan(c0, 0, n_t, 3.0 * float(dp), seed=11)
    R = np.array([10, 20, 30, 100, 140, 180, 240], np.uint8)
    img = np.zeros((1, n_t, 4), np.uint8)
    img[0, :, 0] = R
    img[0, :, 1] = 255 - R
    img[0, :, 2] = (37 * np.arange(n_t)) % 256
    img[0, :, 3] = 255
    uvs = np.zeros((n_t, 3, 2), np.float32)
    uvs[:, :, 0] = ((np.arange(n_t) + 0.5) / n_t)[:, None]   # texel k's centre
    uvs[:, :, 1] = 0.5
    sc = scenes.Scene(meshes=[scenes.Mesh(verts=verts, uvs=uvs, image=img)], N=(N,) * 3,
                      B=(B,) * 3)
    ref = oracle.run(sc)
    col, nrm = ref.mips[0]
    hits, ties = _hit_sets(verts, sc)
    assert target not in ties and hits[target] == list(range(n_t))
    fns = [_face_normal(verts[t].astype(np.float64)) for t in range(n_t)]
    T = img[0].astype(np.float64)
    many = 0
    for vox, h in hits.items():
        if vox in ties:
            continue
        x, y, z = vox
        sl = _slots(h)
        exact = T[sl, :3].sum(axis=0) / len(sl)
        got = col[z, y, x]
        for k in range(3):
            assert _within_one_step(_chan(got, k), exact[k]), (vox, h, k, _chan(got, k), exact[k])
        n = np.sum([fns[t] for t in sl], axis=0)
        n /= np.linalg.norm(n)
        assert np.allclose(nrm[z, y, x], n, atol=2e-5), (vox, h)
        many += len(h) >= 5
    assert many >= 1
    # the target's colour tells the slot rule apart from its neighbours
    got = _chan(col[target[2], target[1], target[0]], 0)
    assert abs(got - R[[0, 1, 2, 6]].mean()) <= 1.0
    for wrong in ([0, 1, 2, 3], list(range(n_t)), [0, 1, 2, 5], [3, 4, 5, 6]):
        assert abs(got - R[wrong].astype(float).mean()) > 4.0


def test_untextured_vertex_normals_signed_dominant_axis():
    # face normal (-0.9, 0.25, 0.35): |x| dominates, but the signed choice of
    # ComputeColorsAsNormals (x > y && x > z, then y > z) projects along z
    n = np.array([-0.9, 0.25, 0.35])
    n /= np.linalg.norm(n)
    u1 = np.cross(n, [0.0, 0.0, 1.0])
    u1 /= np.linalg.norm(u1)
    u2 = np.cross(n, u1)
    c = np.array([0.47, 0.52, 0.49])
    A = c - 0.12 * u1 - 0.10 * u2
    verts = np.array([[A, A + 0.30 * u1, A + 0.28 * u2]], np.float32)
    vn = np.array([[[-0.95, 0.10, 0.30], [-0.70, 0.60, 0.20], [-0.80, -0.10, 0.70]]], np.float32)
    sc = scenes.Scene(meshes=[scenes.Mesh(verts=verts, vnormals=vn)], N=(N,) * 3, B=(B,) * 3)
    ref = oracle.run(sc)
    col, nrm = ref.mips[0]
    offp, dp = _grid_params(sc)
    hits, ties = _hit_sets(verts, sc)
    v = verts[0].astype(np.float64)
    fn = _face_normal(v)
    assert int(np.argmax(np.abs(fn))) == 0
    a, b = _PROJ[2]            # the signed rule's axis
    nA, nB, nC = vn[0].astype(np.float64)
    checked = 0
    differs = 0
    for (x, y, z), h in hits.items():
        if (x, y, z) in ties:
            continue
        ce = _centre(offp, dp, x, y, z)
        u, w = _bary_clamped(ce[[a, b]], v[0][[a, b]], v[1][[a, b]], v[2][[a, b]])
        ni = nA + u * (nC - nA) + w * (nB - nA)
        got = col[z, y, x]
        for k in range(3):   # colour = |interpolated normal| / 1
            assert _within_one_step(_chan(got, k), abs(ni[k]) * 255.0), ((x, y, z), k)
        assert _chan(got, 3) == 255
        assert np.allclose(nrm[z, y, x], ni / np.linalg.norm(ni), atol=2e-5)
        # the absolute-value axis (x: project to (z, y)) would give another value
        ua, wa = _bary_clamped(ce[[2, 1]], v[0][[2, 1]], v[1][[2, 1]], v[2][[2, 1]])
        na = nA + ua * (nC - nA) + wa * (nB - nA)
        differs += np.max(np.abs(np.abs(na) - np.abs(ni))) * 255.0 > 3.0
        checked += 1
    assert checked > 50
    assert differs > checked // 4


def test_untextured_raw_count_division():
    sc0 = scenes.Scene(meshes=[], N=(N,) * 3, B=(B,) * 3)
    offp, dp = _grid_params(sc0)
    target = (15, 17, 16)
    c0 = _centre(offp, dp, *target)
    n_t = 7
    verts = _fan(c0, 0, n_t, 3.0 * float(dp), seed=5)
    sc = scenes.Scene(meshes=[scenes.Mesh(verts=verts)], N=(N,) * 3, B=(B,) * 3)
    ref = oracle.run(sc)
    col, nrm = ref.mips[0]
    hits, ties = _hit_sets(verts, sc)
    assert hits[target] == list(range(n_t))
    fns = [_face_normal(verts[t].astype(np.float64)) for t in range(n_t)]
    far = 0
    for vox, h in hits.items():
        if vox in ties:
            continue
        x, y, z = vox
        sl = _slots(h)
        s = np.sum([fns[t] for t in sl], axis=0) / len(h)     # four slots / raw count
        got = col[z, y, x]
        for k in range(3):
            assert _within_one_step(_chan(got, k), abs(s[k]) * 255.0), (vox, h, k)
        # alpha is min(sum of ones, 1): not divided by the count
        assert _chan(got, 3) == 255
        assert np.allclose(nrm[z, y, x], s / np.linalg.norm(s), atol=2e-5)
        if len(h) >= 5:
            # dividing by the four slots instead would move a channel by more
            # than four steps
            s4 = np.sum([fns[t] for t in sl], axis=0) / len(sl)
            far += np.max(np.abs(np.abs(s4) - np.abs(s))) * 255.0 > 4.0
    assert far >= 1
