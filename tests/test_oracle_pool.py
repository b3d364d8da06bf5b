"""Pins the node-pool / brick-pool oracle (oracle/ref_pool.cpp) on the CPU:
hand-computed textures for small trees, the depth-first (not breadth-first)
block order, the double-precision cube-root sizing at 512 nodes, and the
pointer linkage of random trees checked by walking child pointers."""
import math

import numpy as np
import pytest

import oracle
from tests._gvfiles import brick_payload, node, write_tree

CB = 0x40000000
RB = 0x80000000


def pk(x, y, z):
    return (x << 20) | (y << 10) | z


def origin(b, bx, by):
    return (2 * (b % bx), 2 * ((b // bx) % by), 2 * (b // (bx * by)))


def off(i):
    return (i & 1, (i >> 1) & 1, i >> 2)


def tx(nodes, x, y, z):
    return tuple(int(v) for v in nodes[z, y, x])


def text_trip(c):
    """The reader's value for channel c: '%g' text of c/255, parsed to float,
    times 255.0f, truncated."""
    f = np.float32(float("%g" % (float(np.float32(c)) / 255.0)))
    return int(np.float32(f * np.float32(255.0)))


def test_one_node(tmp_path):
    write_tree(str(tmp_path), node("C", color=(10, 20, 30, 255)), depths=1)
    p = oracle.pool_load(str(tmp_path), slots=(2, 2, 2), store_const=False)
    assert p.node_dims == (2, 2, 2) and p.blocks == 1
    exp = np.zeros((2, 2, 2, 2), np.uint32)
    exp[0, 0, 0, 0] = CB
    assert (p.nodes == exp).all()
    q = oracle.pool_load(str(tmp_path), slots=(2, 2, 2), store_const=True)
    w1 = text_trip(10) | (text_trip(20) << 8) | (text_trip(30) << 16) | (text_trip(255) << 24)
    assert tx(q.nodes, 0, 0, 0) == (CB, w1)
    assert not q.brick_colors.any()


def test_text_round_trip_loses_one_step_for_some_channels():
    lost = [c for c in range(256) if text_trip(c) != c]
    assert all(text_trip(c) == c - 1 for c in lost)
    assert len(lost) == 126


def test_nine_nodes(tmp_path):
    kinds = ["C", "N", "C", "N", "L", "N", "C", "N"]
    kids, b = [], 0
    for k in kinds:
        if k == "N":
            kids.append(node("N", brick=b))
            b += 1
        else:
            kids.append(node(k, color=(1, 2, 3, 4)))
    write_tree(str(tmp_path), node("N", brick=0, children=kids), depths=2, B=4)
    p = oracle.pool_load(str(tmp_path), slots=(2, 2, 2), store_const=False)
    # 16 nodes: floor(16^(1/3)) = 2, x grown to 4
    assert p.node_dims == (4, 2, 2) and p.blocks == 2
    exp = np.zeros((2, 2, 4, 2), np.uint32)
    exp[0, 0, 0] = (RB | pk(2, 0, 0), pk(1, 1, 1))       # root: brick in slot 0
    for i, k in enumerate(kinds):
        x, y, z = 2 + (i & 1), (i >> 1) & 1, i >> 2
        exp[z, y, x] = (CB if k != "N" else 0, 0)
    assert (p.nodes == exp).all()
    col, nrm = brick_payload(0, 0, 6)
    assert (p.brick_colors[:6, :6, :6].ravel() == col).all()
    assert (p.brick_normals[:6, :6, :6].ravel() == nrm).all()
    rest = p.brick_colors.copy()
    rest[:6, :6, :6] = 0
    assert not rest.any()


def test_seventy_three_nodes(tmp_path):
    b1 = 0
    kids = []
    for i in range(8):
        kids.append(node("N", brick=i, children=[node("N", brick=8 * i + j) for j in range(8)]))
    write_tree(str(tmp_path), node("N", brick=0, children=kids), depths=3, B=2)
    p = oracle.pool_load(str(tmp_path), slots=(4, 4, 4), store_const=True)
    # 10 blocks = 80 nodes: floor(80^(1/3)) = 4, x grown 4 -> 6
    assert p.node_dims == (6, 4, 4) and p.blocks == 10
    bx, by = 3, 2
    exp = np.zeros((4, 4, 6, 2), np.uint32)
    exp[0, 0, 0] = (RB | pk(*origin(1, bx, by)), pk(1, 1, 1))
    o1 = origin(1, bx, by)
    for i in range(8):
        x, y, z = (o1[a] + off(i)[a] for a in range(3))
        exp[z, y, x] = (pk(*origin(2 + i, bx, by)), 0)
        oc = origin(2 + i, bx, by)
        for j in range(8):
            gx, gy, gz = (oc[a] + off(j)[a] for a in range(3))
            exp[gz, gy, gx] = (0, 0)
    assert (p.nodes == exp).all()
    assert b1 == 0


def test_depth_first_block_order(tmp_path):
    leaves = lambda k="N", base=0: [node(k, brick=base + j) if k == "N" else node(k)
                                    for j in range(8)]
    g0 = node("N", brick=0, children=leaves("N", 0))
    c0 = node("N", brick=0, children=[g0] + [node("L") for _ in range(7)])
    c2 = node("C", children=[node("L") for _ in range(8)])
    kids = [c0, node("L"), c2] + [node("L") for _ in range(5)]
    write_tree(str(tmp_path), node("N", brick=0, children=kids), depths=4, B=2)
    p = oracle.pool_load(str(tmp_path), slots=(4, 4, 4), store_const=False)
    # depth-first: root -> block 1, c0 -> 2, g0 -> 3, c2 -> 4 (breadth-first
    # would give c2 block 3 and g0 block 4); 40 nodes: 2 -> x grown to 10
    assert p.node_dims == (10, 2, 2) and p.blocks == 5
    bx, by = 5, 1
    n = p.nodes
    assert tx(n, 0, 0, 0) == (RB | pk(*origin(1, bx, by)), pk(1, 1, 1))
    assert tx(n, 2, 0, 0) == (pk(*origin(2, bx, by)), 0)            # c0
    assert tx(n, 3, 0, 0) == (CB, 0)                                 # c1 LEAF-CONST
    assert tx(n, 2, 1, 0) == (CB | pk(*origin(4, bx, by)), 0)        # c2 CONST with kids
    o2 = origin(2, bx, by)
    assert tx(n, *o2) == (pk(*origin(3, bx, by)), 0)                 # g0
    o4 = origin(4, bx, by)
    for j in range(8):
        assert tx(n, *(o4[a] + off(j)[a] for a in range(3))) == (CB, 0)
    o3 = origin(3, bx, by)
    for j in range(8):
        assert tx(n, *(o3[a] + off(j)[a] for a in range(3))) == (0, 0)


def _random_tree(rng, depth, maxd, counters):
    r = rng.random()
    if depth == maxd - 1:
        kind = "N" if r < 0.6 else "L"
    else:
        kind = "N" if r < 0.5 else ("C" if r < 0.7 else "L")
    kids = None
    if depth < maxd - 1 and kind != "L":
        kids = [_random_tree(rng, depth + 1, maxd, counters) for _ in range(8)]
    if kind == "N":
        b = counters[depth]
        counters[depth] += 1
        return node("N", brick=b, children=kids)
    return node(kind, color=tuple(int(c) for c in rng.integers(0, 256, 4)), children=kids)


def _dfs_blocks(root):
    """Depth-first preorder of the nodes with children: block = 1 + rank."""
    order = []

    def walk(n):
        if n["children"]:
            order.append(n)
            for c in n["children"]:
                walk(c)
    walk(root)
    return {id(n): 1 + i for i, n in enumerate(order)}


def _check_links(p, root):
    """Walk child pointers from the root texel and compare with the tree."""
    nd = p.node_dims
    bx, by = nd[0] // 2, nd[1] // 2
    blocks = _dfs_blocks(root)
    seen = set()

    def visit(n, t):
        assert t not in seen
        seen.add(t)
        w0, w1 = tx(p.nodes, *t)
        assert bool(w0 & CB) == (n["kind"] != "N")
        if n["kind"] != "N":
            c = n["color"]
            assert w1 == (text_trip(c[0]) | (text_trip(c[1]) << 8) | (text_trip(c[2]) << 16)
                          | (text_trip(c[3]) << 24))
        ptr = (w0 >> 20 & 1023, w0 >> 10 & 1023, w0 & 1023)
        if n["children"]:
            assert ptr == origin(blocks[id(n)], bx, by)
            for i, c in enumerate(n["children"]):
                visit(c, tuple(ptr[a] + off(i)[a] for a in range(3)))
        else:
            assert ptr == (0, 0, 0)
    visit(root, (0, 0, 0))
    # every other texel is empty
    used = np.zeros(p.nodes.shape[:3], bool)
    for (x, y, z) in seen:
        used[z, y, x] = True
    assert not p.nodes[~used].any()
    assert len(seen) <= 8 * p.blocks


@pytest.mark.parametrize("seed,maxd", [(1, 3), (2, 4), (3, 4), (4, 5)])
def test_random_tree_links(tmp_path, seed, maxd):
    rng = np.random.default_rng(seed)
    counters = [1] + [0] * (maxd - 1)
    root = node("N", brick=0, children=[_random_tree(rng, 1, maxd, counters) for _ in range(8)])
    write_tree(str(tmp_path), root, depths=maxd, B=2)
    p = oracle.pool_load(str(tmp_path), slots=(4, 4, 4), store_const=True)
    nblocks = 1 + len(_dfs_blocks(root))
    assert p.blocks == nblocks
    _check_links(p, root)
    for d in range(maxd):
        fc, fn = p.file_bricks[d]
        assert fc.shape[0] == counters[d]
        for i in range(counters[d]):
            col, nrm = brick_payload(d, i, 4)
            assert (fc[i] == col).all() and (fn[i] == nrm).all()


def test_cube_root_edge_at_512_nodes(tmp_path):
    # 64 blocks = 512 nodes: pow(512, 1/3.0) = 7.999... in double, so the
    # side is 6 (7 made even), x grown to 16 -- not 8^3.
    assert math.floor(math.pow(512, 1.0 / 3.0)) == 7
    kids = []
    n2 = 0
    for i in range(8):
        gk = []
        for j in range(8):
            if n2 < 54:
                gk.append(node("N", brick=n2, children=[node("L") for _ in range(8)]))
                n2 += 1
            else:
                gk.append(node("L"))
        kids.append(node("N", brick=i, children=gk))
    root = node("N", brick=0, children=kids)
    write_tree(str(tmp_path), root, depths=4, B=2)
    p = oracle.pool_load(str(tmp_path), slots=(4, 4, 4), store_const=True)
    assert p.blocks == 64
    assert p.node_dims == (16, 6, 6)
    _check_links(p, root)


def test_rejects_trailing_bytes(tmp_path):
    write_tree(str(tmp_path), node("N", brick=0), depths=1, B=2)
    with open(tmp_path / "voxels_0.gvb", "ab") as f:
        f.write(b"\0")
    with pytest.raises(RuntimeError, match="trailing"):
        oracle.pool_load(str(tmp_path), slots=(2, 2, 2))


def test_rejects_depth_mismatch(tmp_path):
    write_tree(str(tmp_path), node("N", brick=0, children=[node("L") for _ in range(8)]),
               depths=2, B=2)
    txt = (tmp_path / "tree.gvx").read_text().replace('Depth="1"', 'Depth="2"', 1)
    (tmp_path / "tree.gvx").write_text(txt)
    with pytest.raises(RuntimeError, match="Depth"):
        oracle.pool_load(str(tmp_path), slots=(2, 2, 2))
