"""Seeded synthetic triangle scenes for the voxelizer workloads.

The reference voxelizes OpenSceneGraph meshes read from disk
(main.cu Voxelize, 3225-3341); there is no input data in the repository, so
every workload here is generated from a seed:

* fBm heightfield terrain (2-D gradient noise, 6 octaves, lacunarity 2,
  gain 0.5, base frequency 4/E, amplitude 0.2*E around 0.5*E, extent E = 1),
  triangulated cell by cell as (v00, v10, v11), (v00, v11, v01), with
  per-vertex normals from central differences and a 256x1 RGBA8 height-ramp
  texture whose three UVs per triangle sit on one texel centre;
* icosphere "rocks" (subdivision 1 = 80 triangles) resting on the terrain,
  untextured, face normals only;
* uniform random triangle soups with edges of a few voxels.

A mesh is a triangle soup: ``verts`` is float32 (T, 3, 3), ``vnormals`` is
float32 (T, 3, 3) or None, ``uvs`` is float32 (T, 3, 2) or None, ``image`` is
uint8 (h, w, 4) or None, matching cuda::Voxelizer::addTriangleGroup
(Voxelizer.h 128-162).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(state: int):
    """splitmix64 generator (Steele, Lea, Flood 2014)."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def permutation(seed: int) -> np.ndarray:
    """Fisher-Yates shuffle of 0..255 driven by splitmix64(seed), doubled."""
    p = list(range(256))
    g = splitmix64(seed)
    for i in range(255, 0, -1):
        j = next(g) % (i + 1)
        p[i], p[j] = p[j], p[i]
    return np.array(p + p, dtype=np.int64)


_GRAD = np.array([[1, 1], [-1, 1], [1, -1], [-1, -1], [1, 0], [-1, 0], [0, 1], [0, -1]],
                 dtype=np.float64)


def gradient_noise(x: np.ndarray, y: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """2-D gradient (Perlin) noise in float64, quintic fade."""
    xi = np.floor(x).astype(np.int64)
    yi = np.floor(y).astype(np.int64)
    xf = x - xi
    yf = y - yi
    xi &= 255
    yi &= 255

    def grad(ix, iy, dx, dy):
        h = perm[perm[ix] + iy] & 7
        g = _GRAD[h]
        return g[..., 0] * dx + g[..., 1] * dy

    def fade(t):
        return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)

    n00 = grad(xi, yi, xf, yf)
    n10 = grad(xi + 1, yi, xf - 1.0, yf)
    n01 = grad(xi, yi + 1, xf, yf - 1.0)
    n11 = grad(xi + 1, yi + 1, xf - 1.0, yf - 1.0)
    u = fade(xf)
    v = fade(yf)
    nx0 = n00 + u * (n10 - n00)
    nx1 = n01 + u * (n11 - n01)
    return nx0 + v * (nx1 - nx0)


def fbm(x: np.ndarray, y: np.ndarray, seed: int, octaves: int = 6, lacunarity: float = 2.0,
        gain: float = 0.5, base_freq: float = 4.0) -> np.ndarray:
    perm = permutation(seed)
    total = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        total += amp * gradient_noise(x * freq, y * freq, perm)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def ridged(x: np.ndarray, y: np.ndarray, seed: int, octaves: int = 6, lacunarity: float = 2.0,
           gain: float = 0.5, base_freq: float = 4.0) -> np.ndarray:
    """Ridged noise: the normalised sum over octaves of amp * (1 - |noise|)^2
    (the same octaves as fbm, its own permutation)."""
    perm = permutation(seed)
    total = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        r = 1.0 - np.abs(gradient_noise(x * freq, y * freq, perm))
        total += amp * (r * r)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


RAMP_STOPS = [
    (0.00, (30, 90, 30)),
    (0.30, (60, 140, 50)),
    (0.55, (120, 100, 60)),
    (0.80, (130, 130, 130)),
    (1.00, (250, 250, 250)),
]


def height_ramp_texture(width: int = 256) -> np.ndarray:
    """Width x 1 RGBA8 colour ramp with five fixed stops and alpha 255."""
    t = (np.arange(width) + 0.5) / width
    img = np.zeros((1, width, 4), dtype=np.uint8)
    xs = [s[0] for s in RAMP_STOPS]
    for c in range(3):
        ys = [s[1][c] for s in RAMP_STOPS]
        img[0, :, c] = np.clip(np.round(np.interp(t, xs, ys)), 0, 255).astype(np.uint8)
    img[0, :, 3] = 255
    return img


@dataclass
class Mesh:
    verts: np.ndarray
    vnormals: np.ndarray | None = None
    uvs: np.ndarray | None = None
    image: np.ndarray | None = None
    is_terrain: bool = False
    addr: tuple = (1, 1)          # cudaAddressModeClamp on both axes
    name: str = "mesh"

    @property
    def ntri(self) -> int:
        return int(self.verts.shape[0])

    def bbox(self):
        v = self.verts.reshape(-1, 3)
        if v.shape[0] == 0:
            return np.zeros(3, np.float32), np.zeros(3, np.float32)
        return v.min(axis=0).astype(np.float32), v.max(axis=0).astype(np.float32)


@dataclass
class Scene:
    meshes: list
    N: tuple
    B: tuple
    p: tuple = (0.0, 0.0, 0.0)
    delta: tuple = field(default=None)
    name: str = "scene"

    def __post_init__(self):
        if self.delta is None:
            self.delta = tuple(np.float32(1.0) / np.float32(n) for n in self.N)

    @property
    def ntri(self) -> int:
        return sum(m.ntri for m in self.meshes)


def _fbm_heights(g: np.ndarray, seed: int, extent: float) -> np.ndarray:
    """0.5 E + 0.2 E fbm(x / E, y / E) on the grid g x g ([j, i] -> (x_i,
    y_j)) as float32, computed in row blocks on a thread pool (the values
    are elementwise, so the blocking does not change them)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    n1 = g.size
    out = np.empty((n1, n1), np.float32)
    rows = max(1, min(256, (1 << 22) // max(n1, 1)))

    def block(j0):
        j1 = min(j0 + rows, n1)
        X, Y = np.meshgrid(g, g[j0:j1], indexing="xy")
        out[j0:j1] = (0.5 * extent + 0.2 * extent * fbm(X / extent, Y / extent, seed)
                      ).astype(np.float32)

    starts = list(range(0, n1, rows))
    workers = min(len(starts), os.cpu_count() or 1, 16)
    if workers <= 1:
        for j0 in starts:
            block(j0)
    else:
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(block, starts))
    return out


def heightfield_terrain(n_cells: int, seed: int, extent: float = 1.0,
                        tex_width: int = 256) -> Mesh:
    """(n_cells+1)^2 fBm heightfield -> 2*n_cells^2 textured terrain triangles."""
    n = n_cells
    g = np.arange(n + 1, dtype=np.float64) * (extent / n)
    Hf = _fbm_heights(g, seed, extent)
    X, Y = np.meshgrid(g.astype(np.float32), g.astype(np.float32), indexing="xy")
    Xf, Yf = X, Y
    P = np.stack([Xf, Yf, Hf], axis=-1)               # (n+1, n+1, 3)

    # Vertex normals by central differences in fp32 (one-sided at the rim).
    dx = np.float32(extent / n)
    Hp = np.pad(Hf, 1, mode="edge")
    dhdx = (Hp[1:-1, 2:] - Hp[1:-1, :-2]) / (np.float32(2.0) * dx)
    dhdy = (Hp[2:, 1:-1] - Hp[:-2, 1:-1]) / (np.float32(2.0) * dx)
    Nv = np.stack([-dhdx, -dhdy, np.ones_like(dhdx)], axis=-1).astype(np.float32)
    Nv /= np.sqrt((Nv * Nv).sum(axis=-1, keepdims=True)).astype(np.float32)

    v00 = P[:-1, :-1]
    v10 = P[:-1, 1:]
    v11 = P[1:, 1:]
    v01 = P[1:, :-1]
    n00 = Nv[:-1, :-1]
    n10 = Nv[:-1, 1:]
    n11 = Nv[1:, 1:]
    n01 = Nv[1:, :-1]
    t1 = np.stack([v00, v10, v11], axis=2).reshape(-1, 3, 3)
    t2 = np.stack([v00, v11, v01], axis=2).reshape(-1, 3, 3)
    m1 = np.stack([n00, n10, n11], axis=2).reshape(-1, 3, 3)
    m2 = np.stack([n00, n11, n01], axis=2).reshape(-1, 3, 3)
    # interleave per cell: (cell tri 1, cell tri 2)
    verts = np.stack([t1, t2], axis=1).reshape(-1, 3, 3).astype(np.float32)
    vnorm = np.stack([m1, m2], axis=1).reshape(-1, 3, 3).astype(np.float32)

    hmean = verts[:, :, 2].astype(np.float64).mean(axis=1)
    lo, hi = float(Hf.min()), float(Hf.max())
    t = (hmean - lo) / max(hi - lo, 1e-12)
    idx = np.clip((t * tex_width).astype(np.int64), 0, tex_width - 1)
    u = ((idx.astype(np.float32) + np.float32(0.5)) / np.float32(tex_width)).astype(np.float32)
    uvs = np.zeros((verts.shape[0], 3, 2), dtype=np.float32)
    uvs[:, :, 0] = u[:, None]
    uvs[:, :, 1] = np.float32(0.5)
    return Mesh(verts=verts, vnormals=vnorm, uvs=uvs, image=height_ramp_texture(tex_width),
                is_terrain=True, name=f"terrain{n + 1}")


def _icosphere1():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(a, dtype=np.float64) / np.linalg.norm(a) for a in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    tris = []
    for a, b, c in f:
        A, Bv, C = v[a], v[b], v[c]
        ab = (A + Bv) / np.linalg.norm(A + Bv)
        bc = (Bv + C) / np.linalg.norm(Bv + C)
        ca = (C + A) / np.linalg.norm(C + A)
        tris += [(A, ab, ca), (Bv, bc, ab), (C, ca, bc), (ab, bc, ca)]
    return np.array(tris, dtype=np.float64)        # (80, 3, 3) unit sphere


def rocks_on_terrain(terrain_heights_fn, count: int, seed: int, rmin: float = 0.001,
                     rmax: float = 0.02, extent: float = 1.0) -> Mesh:
    rng = np.random.Generator(np.random.PCG64(seed))
    sphere = _icosphere1()
    r = np.exp(rng.uniform(np.log(rmin * extent), np.log(rmax * extent), size=count))
    cx = rng.uniform(0.02 * extent, 0.98 * extent, size=count)
    cy = rng.uniform(0.02 * extent, 0.98 * extent, size=count)
    cz = terrain_heights_fn(cx, cy) + 0.3 * r
    c = np.stack([cx, cy, cz], axis=-1)
    verts = (sphere[None, :, :, :] * r[:, None, None, None] + c[:, None, None, :])
    return Mesh(verts=verts.reshape(-1, 3, 3).astype(np.float32), is_terrain=False,
                name=f"rocks{count}")


def terrain_height_fn(seed: int, extent: float = 1.0):
    def h(x, y):
        return 0.5 * extent + 0.2 * extent * fbm(np.asarray(x) / extent, np.asarray(y) / extent,
                                                 seed)
    return h


def random_soup(ntri: int, seed: int, N: int, edge_voxels=(0.5, 4.0), extent: float = 1.0,
                textured: bool = False, uv_mode: str = "centre", vertex_normals: bool = False,
                translucent: bool = False, addr=(1, 1)) -> Mesh:
    """Uniform random triangle soup inside the unit cube, edges of a few voxels.

    uv_mode "centre": the three UVs of a triangle sit on one texel centre;
    "random": independent UVs in [-0.5, 1.5] (bilinear weights, address modes).
    translucent: alpha drawn from {0, 1, 128, 255} (the 0.005 cut-off)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vox = extent / N
    c = rng.uniform(0.05 * extent, 0.95 * extent, size=(ntri, 1, 3))
    scale = rng.uniform(edge_voxels[0], edge_voxels[1], size=(ntri, 1, 1)) * vox
    d = rng.normal(size=(ntri, 3, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    verts = (c + d * scale).astype(np.float32)
    m = Mesh(verts=verts, is_terrain=False, name=f"soup{ntri}", addr=tuple(addr))
    if vertex_normals:
        vn = rng.normal(size=(ntri, 3, 3))
        vn /= np.linalg.norm(vn, axis=-1, keepdims=True)
        m.vnormals = vn.astype(np.float32)
    if textured:
        img = rng.integers(0, 256, size=(4, 4, 4), dtype=np.uint8)
        if translucent:
            img[..., 3] = rng.choice(np.array([0, 1, 128, 255], dtype=np.uint8), size=(4, 4))
        else:
            img[..., 3] = 255
        if uv_mode == "random":
            m.uvs = rng.uniform(-0.5, 1.5, size=(ntri, 3, 2)).astype(np.float32)
        else:
            ij = rng.integers(0, 4, size=(ntri, 2))
            uv = ((ij.astype(np.float32) + np.float32(0.5)) / np.float32(4.0)).astype(np.float32)
            m.uvs = np.repeat(uv[:, None, :], 3, axis=1).astype(np.float32)
        m.image = img
    return m


def scene_small_terrain(seed: int = 1) -> Scene:
    """129x129 fBm heightfield (32,768 triangles), 128^3 grid of 8^3 bricks."""
    return Scene(meshes=[heightfield_terrain(128, seed)], N=(128, 128, 128), B=(8, 8, 8),
                 name="terrain129_N128_B8")


def scene_full_octree(seed: int = 2, rock_seed: int = 3, n_rocks: int = 2000,
                      cells: int = 1024, N: int = 1024, B: int = 16) -> Scene:
    """1025^2 fBm heightfield (~2.1M triangles) + 2,000 icosphere rocks,
    1024^3 grid of 16^3 bricks, octree levels 0-6."""
    terrain = heightfield_terrain(cells, seed)
    rocks = rocks_on_terrain(terrain_height_fn(seed), n_rocks, rock_seed)
    return Scene(meshes=[terrain, rocks], N=(N, N, N), B=(B, B, B),
                 name=f"terrain{cells + 1}_rocks{n_rocks}_N{N}_B{B}")


def scene_two_triangles() -> Scene:
    """The two-triangle case of main.cu 4744-4781 (64^3 grid, 8^3 bricks, p=0,
    delta=1), with flat face normals and no texture."""
    v = np.array([[[0.5, 15.5, 7.5], [7.5, 15.5, 7.5], [7.5, 23.5, 7.5]],
                  [[2.5, 0.5, 7.5], [2.5, 2.5, 7.5], [0.5, 2.5, 7.5]]], dtype=np.float32)
    return Scene(meshes=[Mesh(verts=v, name="two_tris")], N=(64, 64, 64), B=(8, 8, 8),
                 p=(0.0, 0.0, 0.0), delta=(1.0, 1.0, 1.0), name="two_triangles")


def scene_mixed(seed: int = 7, N: int = 64, B: int = 8, order: str = "terrain_first") -> Scene:
    """Several groups in one grid: textured terrain (fill), untextured rocks,
    a textured soup with vertex normals, random UVs and wrap addressing, and
    an untextured soup with vertex normals.  order "rocks_first" puts the rocks
    before the terrain (the first group is not a terrain, the fill is not the
    first group's); "two_terrains" replaces the textured soup by a second
    terrain (another seed) after the rocks, whose fill lands on voxels the
    first terrain's fill and the rocks already hold."""
    cells = N // 2
    terrain = heightfield_terrain(cells, seed)
    rocks = rocks_on_terrain(terrain_height_fn(seed), 40, seed + 1, rmin=0.01, rmax=0.06)
    soup_t = random_soup(600, seed + 2, N, textured=True, uv_mode="random", vertex_normals=True,
                         translucent=True, addr=(0, 2))
    soup_u = random_soup(400, seed + 3, N, vertex_normals=True)
    if order == "rocks_first":
        meshes = [rocks, terrain, soup_t, soup_u]
    elif order == "two_terrains":
        meshes = [terrain, rocks, heightfield_terrain(cells, seed + 20), soup_u]
    else:
        meshes = [terrain, rocks, soup_t, soup_u]
    suffix = "" if order == "terrain_first" else "_" + order
    return Scene(meshes=meshes, N=(N, N, N), B=(B, B, B), name=f"mixed_N{N}_B{B}{suffix}")


def scene_many_groups(seed: int = 100, N: int = 64, B: int = 8) -> Scene:
    """Twelve groups in one grid: nine random soups (every third textured
    with random UVs, every second with vertex normals), a soup wholly
    outside the grid (no bins), an empty group and a textured terrain as the
    eleventh group, whose fill lands on what ten groups before it hold."""
    meshes = [random_soup(120 + 40 * i, seed + i, N, textured=(i % 3 == 0), uv_mode="random",
                          vertex_normals=(i % 2 == 1)) for i in range(9)]
    far = random_soup(50, seed + 100, N)
    far.verts = (far.verts + np.float32(3.0)).astype(np.float32)
    meshes.insert(3, far)
    meshes.insert(6, Mesh(verts=np.zeros((0, 3, 3), np.float32), name="empty"))
    meshes.insert(10, heightfield_terrain(N // 2, seed + 17))
    return Scene(meshes=meshes, N=(N, N, N), B=(B, B, B), name=f"many_groups_N{N}_B{B}")


def scene_big_triangles(seed: int = 13, N: int = 128, B: int = 8) -> Scene:
    """One untextured group mixing small random triangles with two that span
    the whole grid (18^3 tiles at N=128, B=8: more (triangle, tile) bins than
    the binning lists per triangle, so those two walk their boxes instead),
    each placed among the small ones of a different workgroup of 256."""
    big = np.array([[[-0.02, -0.02, -0.01], [1.02, 0.01, 1.01], [0.0, 1.02, 1.0]],
                    [[1.02, 1.02, -0.02], [-0.02, 0.99, 0.5], [0.98, -0.02, 1.02]]],
                   dtype=np.float32)
    soup = random_soup(500, seed, N)
    verts = np.concatenate([soup.verts[:250], big[:1], soup.verts[250:], big[1:]])
    return Scene(meshes=[Mesh(verts=verts.astype(np.float32), name="big_and_small")],
                 N=(N, N, N), B=(B, B, B), name=f"big_tris_N{N}_B{B}")


def scene_terrain_n(cells: int, N: int, B: int, seed: int = 11) -> Scene:
    """Heightfield of (cells+1)^2 vertices on an N^3 grid of B^3 bricks."""
    return Scene(meshes=[heightfield_terrain(cells, seed)], N=(N, N, N), B=(B, B, B),
                 name=f"terrain{cells + 1}_N{N}_B{B}")


def flythrough(seed: int = 5, frames: int = 300, n_points: int = 12, extent: float = 1.0,
               terrain_seed: int = 2, alt=(0.05, 0.2), height_fn=None):
    """Seeded Catmull-Rom camera path over the terrain: n_points control
    points uniform in the inner 80% of the xy extent at altitude alt above the
    fBm surface (or height_fn's); the camera looks along the path's
    horizontal heading, tilted down by atan(0.1 / 0.25).  Returns a list of
    (position, look, up)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = height_fn if height_fn is not None else terrain_height_fn(terrain_seed, extent)
    xy = rng.uniform(0.1 * extent, 0.9 * extent, size=(n_points, 2))
    z = h(xy[:, 0], xy[:, 1]) + rng.uniform(alt[0] * extent, alt[1] * extent, size=n_points)
    P = np.concatenate([xy, z[:, None]], axis=1)
    P = np.concatenate([P[:1], P, P[-1:]], axis=0)   # end tangents

    def cr(t):
        seg = min(int(t), n_points - 2)
        u = t - seg
        p0, p1, p2, p3 = P[seg], P[seg + 1], P[seg + 2], P[seg + 3]
        return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u * u
                      + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)

    out = []
    span = n_points - 1
    for f in range(frames):
        t = span * f / max(frames - 1, 1)
        pos = cr(t)
        if t + 0.25 <= span:
            ahead = cr(t + 0.25)
        else:                                    # past the last point: keep heading
            ahead = pos + (pos - cr(max(t - 0.25, 0.0)))
        fwd = ahead - pos
        fwd[2] = 0.0
        if np.linalg.norm(fwd) < 1e-9:
            fwd = np.array([1.0, 0.0, 0.0])
        fwd = 0.25 * extent * fwd / np.linalg.norm(fwd)
        look = fwd + np.array([0.0, 0.0, -0.1 * extent])
        look = look / np.linalg.norm(look)
        up = np.array([0.0, 0.0, 1.0]) - look[2] * look
        up /= np.linalg.norm(up)
        out.append((pos.astype(np.float32), look.astype(np.float32), up.astype(np.float32)))
    return out


def batch_scenes(seed: int = 4, n_jobs: int = 256, tmin: float = 2e3, tmax: float = 2e6,
                 grids=(64, 128, 256, 512, 1024)):
    """Independent meshes for batched voxelization: triangle counts
    log-uniform in [tmin, tmax], half fBm heightfield tiles (2*cells^2
    triangles, cells = sqrt(T/2) rounded to a power of two between 16 and
    1024) and half uniform random soups with edges of 0.5-4 voxels; each on
    a grid drawn from `grids`, at most 8x the mesh's resolution."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for j in range(n_jobs):
        T = float(np.exp(rng.uniform(np.log(tmin), np.log(tmax))))
        N = int(rng.choice(grids))
        if j % 2 == 0:
            cells = int(2 ** np.clip(np.round(np.log2(np.sqrt(T / 2.0))), 4, 10))
            mesh = heightfield_terrain(cells, seed=seed * 1000 + j)
            N = min(N, max(64, 8 * cells))
            out.append(Scene(meshes=[mesh], N=(N, N, N), B=(8, 8, 8),
                             name=f"job{j}_terrain{cells + 1}_N{N}"))
        else:
            mesh = random_soup(int(T), seed * 1000 + j, N)
            out.append(Scene(meshes=[mesh], N=(N, N, N), B=(8, 8, 8),
                             name=f"job{j}_soup{int(T)}_N{N}"))
    return out
