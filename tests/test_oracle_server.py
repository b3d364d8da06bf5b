"""Pins the brick-server oracle (oracle/server.py) on a nine-node tree with a
two-slot pool, frame by frame: FIFO order within a frame, eviction of the
least recently used slot (ties by slot index), protection of slots used in
the current frame, the parent rule, deferral when no slot is free, and the
node texel bits after each admission."""
import numpy as np

from oracle.server import NONE, RB, RefServer

CB = 0x40000000


def pk(x, y, z):
    return (x << 20) | (y << 10) | z


def tree():
    tex = np.zeros((2, 2, 4, 2), np.uint32)            # 16 texels: dims (4, 2, 2)
    tex[0, 0, 0] = (RB | pk(2, 0, 0), pk(1, 1, 1))     # root: brick in slot 0
    types1 = np.array([1, 1, 1, 0, 1, 2, 1, 1], np.uint8)
    texel1 = np.zeros(8, np.uint32)
    for i in range(8):
        x, y, z = 2 + (i & 1), (i >> 1) & 1, i >> 2
        tex[z, y, x, 0] = CB if types1[i] != 1 else 0
        texel1[i] = (z * 2 + y) * 4 + x
    types = [np.array([1], np.uint8), types1]
    bidx = [np.array([0], np.uint32), np.cumsum(types1 == 1).astype(np.uint32) - 1]
    texel = [np.array([0], np.uint32), texel1]
    return tex, types, bidx, texel


def nid(t, nd=(4, 2, 2)):
    x, y, z = t % nd[0], (t // nd[0]) % nd[1], t // (nd[0] * nd[1])
    return ((x + 1) << 20) | ((y + 1) << 10) | (z + 1)


def test_lru_fifo_parent_and_deferral():
    tex, types, bidx, texel = tree()
    s = RefServer(tex, types, bidx, texel, slots=(2, 1, 1), B=(2, 2, 2), max_up=2)
    c = [int(t) for t in texel[1]]
    assert s.node_slot == {0: 0}
    # frame 1: two children; slot 1 is empty (key 0), slot 0 holds the root
    # (unused this frame): both taken, in slot order, FIFO in texel order
    assert s.step([nid(c[1]), nid(c[0]), nid(c[3])]) == 2       # c3 is CONST: ignored
    assert s.admitted[-1] == [(c[0], 0), (c[1], 1)]
    t = s.texture().reshape(-1, 2)
    assert not (t[0, 0] & RB)                                    # root evicted
    assert t[c[0], 0] & RB and t[c[0], 1] == pk(1, 1, 1)
    assert t[c[1], 0] & RB and t[c[1], 1] == pk(5, 1, 1)         # slot 1 origin x = 4
    # frame 2: c0 used again (slot 0 protected); c2 requested: its parent,
    # the root, is NON-CONST without a brick and is queued before it
    assert s.step([nid(c[0]), nid(c[2])]) == 1
    assert s.admitted[-1] == [(0, 1)]                            # root into slot 1
    assert list(s.fifo) == [c[2]]                                # deferred
    t = s.texture().reshape(-1, 2)
    assert not (t[c[1], 0] & RB) and t[0, 0] & RB
    # frame 3: nothing requested; both slots age 1, the lower index goes
    assert s.step([]) == 1
    assert s.admitted[-1] == [(c[2], 0)]
    assert s.node_slot == {0: 1, c[2]: 0}
    assert not s.fifo and not s.pending


def test_requests_outside_tree_and_constants_are_ignored():
    tex, types, bidx, texel = tree()
    s = RefServer(tex, types, bidx, texel, slots=(2, 1, 1), B=(2, 2, 2), max_up=2)
    c = [int(t) for t in texel[1]]
    bad = nid(1)                     # a sibling texel of the root: no node there
    assert 1 not in [int(t) for t in texel[1]] + [0]
    assert s.step([bad, nid(c[3]), nid(c[5])]) == 0    # CONST, LEAF-CONST
    assert not s.fifo
    # the resident root touched: its slot is protected this frame
    assert s.step([nid(0), nid(c[0]), nid(c[1])]) == 1
    assert s.admitted[-1] == [(c[0], 1)]
    assert list(s.fifo) == [c[1]]


def _run_both(requests):
    """The single-tree oracle and the shared-pool oracle with that one tree
    attached: the same admissions, textures and slot tables frame by frame."""
    from oracle.server import RefSharedServer
    tex, types, bidx, texel = tree()
    a = RefServer(tex, types, bidx, texel, slots=(2, 1, 1), B=(2, 2, 2), max_up=2)
    b = RefSharedServer(slots=(2, 1, 1), B=(2, 2, 2), max_trees=3, cap=16, max_up=2)
    assert b.attach(tex, types, bidx, texel) == 0
    for ids in requests:
        assert a.step(ids) == b.step([(0, ids)])
        assert a.admitted[-1] == b.admitted[-1]
        assert (a.texture() == b.texture(0)).all()
        assert a.node_slot == b.node_slot
        assert list(a.fifo) == list(b.fifo)


def test_shared_oracle_with_one_tree_is_the_single_tree_oracle():
    tex, types, bidx, texel = tree()
    c = [int(t) for t in texel[1]]
    _run_both([[nid(c[1]), nid(c[0]), nid(c[3])], [nid(c[0]), nid(c[2])], [],
               [nid(1), nid(c[5]), nid(c[6])], [nid(0), nid(c[7]), nid(c[4])], []])


def _no_root_tree():
    tex, types, bidx, texel = tree()
    tex = tex.copy()
    tex[0, 0, 0] = (pk(2, 0, 0), 0)      # root brick not uploaded
    return tex, types, bidx, texel


def test_shared_attach_seeds_root_and_detach_frees_slots():
    from oracle.server import RefSharedServer
    s = RefSharedServer(slots=(2, 1, 1), B=(2, 2, 2), max_trees=2, cap=16, max_up=2)
    A = tree()
    Bt = _no_root_tree()
    c = [int(t) for t in A[3][1]]
    assert s.attach(*A) == 0 and s.node_slot == {0: 0}          # root adopted
    assert s.attach(*Bt) == 1 and s.pending == {16}             # root pending
    assert s.step([]) == 1 and s.admitted[-1] == [(16, 1)]      # empty slot first
    # A's c0 (parent resident) and B's c1 (parent resident): both slots age
    # out, the older (A's root, key 254) first
    assert s.step([(0, [nid(c[0])]), (1, [nid(c[1])])]) == 2
    assert s.admitted[-1] == [(c[0], 0), (16 + c[1], 1)]
    assert not (s.texture(0).reshape(-1, 2)[0, 0] & RB)         # A root evicted
    assert not (s.texture(1).reshape(-1, 2)[0, 0] & RB)         # B root evicted
    s.detach(0)
    assert s.slot_node[0] is None and s.node_slot == {16 + c[1]: 1}
    # B's c2: its parent (B's root) lost its brick, both queued, root first
    assert s.step([(1, [nid(c[2])])]) == 2
    assert s.admitted[-1] == [(16, 0), (16 + c[2], 1)]
    t1 = s.texture(1).reshape(-1, 2)
    assert t1[0, 0] & RB and t1[0, 1] == pk(1, 1, 1)
    assert not (t1[c[1], 0] & RB) and t1[c[2], 0] & RB and t1[c[2], 1] == pk(5, 1, 1)


def test_shared_detach_drops_queued_entries():
    from oracle.server import RefSharedServer
    s = RefSharedServer(slots=(2, 1, 1), B=(2, 2, 2), max_trees=2, cap=16, max_up=1)
    A = tree()
    c = [int(t) for t in A[3][1]]
    s.attach(*A)
    s.attach(*_no_root_tree())
    assert s.step([]) == 1 and s.admitted[-1] == [(16, 1)]
    assert s.step([(0, [nid(c[0]), nid(c[1])])]) == 1
    assert s.admitted[-1] == [(c[0], 0)] and list(s.fifo) == [c[1]]
    s.detach(0)
    assert not s.fifo and not s.pending and s.slot_node == [None, 16]
    assert s.step([]) == 0 and s.admitted[-1] == []
    # B's c0 and c1 queued in texel order; one admission, into the empty slot
    assert s.step([(1, [nid(c[1]), nid(c[0])])]) == 1
    assert s.admitted[-1] == [(16 + c[0], 0)] and list(s.fifo) == [16 + c[1]]


def test_wide_pool_same_policy_slot_in_word_one():
    """A wide pool (slots (1, 1, n), word 1 of a resident node its slot)
    admits the same (texel, slot) pairs as the reference's pool of n slots;
    only word 1 changes: the slot instead of its texel origin + 1."""
    tex, types, bidx, texel = tree()
    wtex = tex.copy()
    wtex[0, 0, 0, 1] = 0                                       # root's brick: slot 0
    ref = RefServer(tex, types, bidx, texel, slots=(2, 1, 1), B=(2, 2, 2), max_up=2)
    wide = RefServer(wtex, types, bidx, texel, slots=(1, 1, 2), B=(2, 2, 2), max_up=2, wide=True)
    assert wide.node_slot == ref.node_slot == {0: 0}
    c = [int(t) for t in texel[1]]
    for ids in ([nid(c[1]), nid(c[0]), nid(c[3])], [nid(c[0]), nid(c[2])], [], [nid(c[6])]):
        assert wide.step(ids) == ref.step(ids)
        assert wide.admitted[-1] == ref.admitted[-1]
        a, b = wide.texture().reshape(-1, 2), ref.texture().reshape(-1, 2)
        assert (a[:, 0] == b[:, 0]).all()
        for t, s in wide.node_slot.items():
            assert a[t, 1] == s and b[t, 1] == pk(1 + 4 * s, 1, 1)
