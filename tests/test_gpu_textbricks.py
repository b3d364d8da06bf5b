"""Text bricks (GPU): generateVoxelsAndOctTree(outputDir, outputBinary=False)
writes voxels_<depth>.gvt as the reference's WriteTextBrick does
(VoxelBrickWriter.cu 207-325, 45-68, 96-103): per brick in index order a
<Brick> line with the window's offset in its mip level (0 for a brick that
reached the far side of the chunk and was completed from its own partial
buffer), dimensions and borders, then the window's colours as raw bytes and
its float normals in default iostream formatting.  The expected files are
restated here from the oracle's mip levels and brick indices and compared
byte for byte; tree.gvx says Binary="NO"."""
import numpy as np
import pytest

import oracle
from voxvizthesis_amd import scenes
from voxvizthesis_amd.voxelizer import Voxelizer

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def expected_gvt(ref, B, extra, chunk=None):
    """chunk: the chunk edge at level 0 with several chunks (the levels come
    back chunk by chunk, side by side): a brick whose window leaves one chunk
    of chunk >> k voxels is partial, the others are labelled with their
    offset in their chunk."""
    L = ref.levels
    bord = [list(extra)]
    for k in range(1, L):
        bord.append([max(b >> 1, 1) for b in bord[-1]])
    files = []
    for d in range(L):
        k = L - 1 - d
        col, nrm = ref.mips[k]
        col = np.asarray(col, np.uint32)
        nrm = np.asarray(nrm, np.float32)
        S = [col.shape[2], col.shape[1], col.shape[0]]
        bidx = np.asarray(ref.tree[d]["bidx"], np.uint64)
        nodes = np.nonzero(bidx != NONE)[0]
        nodes = nodes[np.argsort(bidx[nodes], kind="stable")]
        R = 1 << d
        out = [b"<Bricks>\n"]
        for n in nodes:
            n = int(n)
            c = (n % R, (n // R) % R, n // (R * R))
            m = [c[a] * B + bord[k][a] for a in range(3)]
            x0 = [v - 1 for v in m]
            if chunk is None:
                partial = any(m[a] + B >= S[a] for a in range(3))
                w0 = [0, 0, 0] if partial else x0
            else:
                ck = chunk >> k
                partial = any(x0[a] // ck != (x0[a] + B + 1) // ck for a in range(3))
                w0 = [0, 0, 0] if partial else [x0[a] - (x0[a] // ck) * ck for a in range(3)]
            e = B + 2
            out.append((f' <Brick X="{w0[0]}" Y="{w0[1]}" Z="{w0[2]}" DimX="{e}" DimY="{e}" '
                        f'DimZ="{e}" BorderX="1" BorderY="1" BorderZ="1" />\n').encode())
            out.append(b"  <BrickColors>\n")
            for z in range(e):
                for y in range(e):
                    out.append(f'    <Row Y="{w0[1] + y}" Z="{w0[2] + z}">\n'.encode())
                    row = col[x0[2] + z, x0[1] + y, x0[0]:x0[0] + e]
                    for v in row:
                        v = int(v)
                        out.append(b"(" + bytes([v & 255]) + b"," + bytes([(v >> 8) & 255]) + b","
                                   + bytes([(v >> 16) & 255]) + b"," + bytes([v >> 24]) + b") ")
                    out.append(b"\n    </Row>\n")
            out.append(b"  </BrickColors>\n")
            out.append(b"  <BrickGradients>\n")
            for z in range(e):
                for y in range(e):
                    out.append(f'    <Row YZ="{w0[1] + y} {w0[2] + z}">\n'.encode())
                    row = nrm[x0[2] + z, x0[1] + y, x0[0]:x0[0] + e]
                    for v in row:
                        out.append(("(%g,%g,%g) " % (float(v[0]), float(v[1]), float(v[2])))
                                   .encode())
                    out.append(b"\n    </Row>\n")
            out.append(b"  </BrickGradients>\n")
            out.append(b" </Brick>\n")
        out.append(b"</Bricks>\n")
        files.append(b"".join(out))
    return files


@pytest.mark.parametrize("name", ["two_triangles", "small_terrain", "mixed"])
def test_text_bricks_equal_restatement(name, tmp_path):
    sc = {"two_triangles": scenes.scene_two_triangles,
          "small_terrain": scenes.scene_small_terrain,
          "mixed": scenes.scene_mixed}[name]()
    print(f"[gvt] {sc.name}", flush=True)
    ref = oracle.run(sc, keep_mips=True)
    vz = Voxelizer(sc.N, sc.B)
    try:
        ok = vz.initCuda() and vz.allocateVoxelMemory() and vz.setVoxelizationParams(sc.p,
                                                                                  sc.delta)
        for m in sc.meshes:
            ok = ok and vz.addMesh(m)
        assert ok and vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()
        assert vz.generateVoxelsAndOctTree(str(tmp_path), False, False) == ref.status, \
            vz.getErrorMessage()
        B = sc.B[0]
        extra = [(n // b) >> 1 for n, b in zip(sc.N, sc.B)]
        want = expected_gvt(ref, B, extra)
        nbricks = 0
        for d, w in enumerate(want):
            got = (tmp_path / f"voxels_{d}.gvt").read_bytes()
            assert got == w, f"depth {d}"
            nbricks += w.count(b"<Brick ")
        assert nbricks > 0
        tree = (tmp_path / "tree.gvx").read_text()
        assert 'Binary="NO"' in tree and 'Compressed="NO"' in tree
        print(f"[gvt] {nbricks} bricks equal", flush=True)
    finally:
        vz.close()


def _brick_colours(gvt: bytes, e: int):
    """The colour bytes of every brick of a .gvt file, (nbricks, e^3) uint32."""
    out = []
    for part in gvt.split(b"  <BrickColors>\n")[1:]:
        body = part.split(b"  </BrickColors>")[0]
        vals = []
        # each voxel is "(" r "," g "," b "," a ") " with raw bytes: ten
        # bytes from its "(" (rows and their headers hold no other "(")
        i = 0
        while True:
            j = body.find(b"(", i)
            if j < 0:
                break
            r, g, b, a = body[j + 1], body[j + 3], body[j + 5], body[j + 7]
            vals.append(r | (g << 8) | (b << 16) | (a << 24))
            i = j + 10
        assert len(vals) == e ** 3
        out.append(vals)
    return np.asarray(out, np.uint32).reshape(-1, e ** 3)


@pytest.mark.parametrize("name,chunk", [("mixed", 8), ("mixed", 16), ("terrain", 16)])
def test_text_bricks_several_chunks(name, chunk, tmp_path):
    """With several chunks: the files equal the restatement from the
    oracle's chunked run (levels chunk by chunk, partial bricks labelled 0,
    the others with their offset in their chunk), and every brick's colours
    equal the oracle's assembled brick of the same index."""
    sc = {"mixed": lambda: scenes.scene_mixed(seed=3, N=32, B=4),
          "terrain": lambda: scenes.scene_terrain_n(32, 64, 8, seed=4)}[name]()
    print(f"[gvt] {sc.name} chunk {chunk}", flush=True)
    c3 = (chunk, chunk, chunk)
    ref = oracle.run(sc, keep_mips=True, chunk=c3)
    vz = Voxelizer(sc.N, sc.B, c3)
    try:
        ok = vz.initCuda() and vz.allocateVoxelMemory() and vz.setVoxelizationParams(sc.p,
                                                                                  sc.delta)
        for m in sc.meshes:
            ok = ok and vz.addMesh(m)
        assert ok and vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()
        assert vz.generateVoxelsAndOctTree(str(tmp_path), False, False) == ref.status, \
            vz.getErrorMessage()
        B = sc.B[0]
        extra = [(n // b) >> 1 for n, b in zip(sc.N, sc.B)]
        want = expected_gvt(ref, B, extra, chunk)
        nbricks = partials = 0
        for d, w in enumerate(want):
            got = (tmp_path / f"voxels_{d}.gvt").read_bytes()
            assert got == w, f"depth {d}"
            nbricks += w.count(b"<Brick ")
            partials += w.count(b'<Brick X="0" Y="0" Z="0"')
            if ref.tree[d]["nb"]:
                assert np.array_equal(_brick_colours(got, B + 2), ref.tree[d]["bcol"]), d
        assert partials > 0
        # 8-voxel chunks leave no 6-voxel brick window inside one chunk;
        # 16-voxel chunks hold some
        assert (nbricks > partials) == (chunk >= 16)
        print(f"[gvt] {nbricks} bricks equal, {partials} partial", flush=True)
    finally:
        vz.close()
