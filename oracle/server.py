"""CPU restatement of the brick server's per-frame policy (small cases).

TEST INFRASTRUCTURE ONLY: the oracle the device brick server is checked
against; only tests/ import it.

The reference's policy (GigaVoxelsOctTree.cpp 430-652, GigaVoxelsBrickPool.cpp
406-427, 561-682): every used node with a resident brick moves its slot to the
back of an LRU list; a NON-CONST node without a brick goes on the upload list
once; BrickPool::update copies each listed brick into the front (least
recently used) slot, evicting its node, and the node's NON-CONST parent
without a brick too.  Restated deterministically, frame by frame:
  touch      slots of requested resident nodes get the frame number;
  pending    requested NON-CONST nodes without a brick, and their NON-CONST
             parents without one, join a FIFO once (ascending texel order
             within a frame);
  admission  min(FIFO length, uploads per frame, slots not used this frame)
             entries take the slots with key (empty 0, else 256 - min(age,
             255)) smallest, ties by slot index, assigned in ascending slot
             order; an evicted node loses its resident bit.
"""
from __future__ import annotations

from collections import deque

import numpy as np

NONE = 0xFFFFFFFF
RB = 0x80000000


class RefServer:
    def __init__(self, tex, types, bidx, texel, slots, B, max_up, wide=False):
        """tex: (z, y, x, 2) node texture after the octree is built; types /
        bidx / texel: per-depth arrays of the tree; slots: pool slots per axis
        (wide: a wide pool, slots (1, 1, n), word 1 of a resident node its slot)."""
        self.wide = wide
        self.nd = (tex.shape[2], tex.shape[1], tex.shape[0])
        self.tex = tex.reshape(-1, 2).copy()
        self.slots = tuple(slots)
        self.e = tuple(b + 2 for b in B)
        self.nslots = slots[0] * slots[1] * slots[2]
        self.max_up = max_up if max_up else self.nslots
        self.types, self.bidx = types, bidx
        self.info = {}
        for d, tx in enumerate(texel):
            for n in np.nonzero(tx != NONE)[0]:
                self.info[int(tx[n])] = (d, int(n))
        self.texel = texel
        self.node_slot = {}
        self.slot_node = [None] * self.nslots
        self.slot_last = [0] * self.nslots
        for t in np.nonzero(self.tex[:, 0] & RB)[0]:
            w1 = int(self.tex[t, 1])
            x = (((w1 >> 20) & 1023) - 1) // self.e[0]
            y = (((w1 >> 10) & 1023) - 1) // self.e[1]
            z = ((w1 & 1023) - 1) // self.e[2]
            s = w1 if wide else (z * slots[1] + y) * slots[0] + x
            self.node_slot[int(t)] = s
            self.slot_node[s] = int(t)
        self.fifo = deque()
        self.pending = set()
        self.frame = 0
        self.admitted = []          # per frame: list of (texel, slot)

    def _non_const(self, t):
        d, n = self.info[t]
        return int(self.types[d][n]) == 1

    def _parent(self, t):
        d, n = self.info[t]
        if d == 0:
            return None
        R = 1 << d
        x, y, z = n % R, (n // R) % R, n // (R * R)
        p = ((z >> 1) * (R >> 1) + (y >> 1)) * (R >> 1) + (x >> 1)
        v = int(self.texel[d - 1][p])
        return None if v == NONE else v

    def _slot_origin(self, s):
        sx, sy = self.slots[0], self.slots[1]
        return ((s % sx) * self.e[0], ((s // sx) % sy) * self.e[1], (s // (sx * sy)) * self.e[2])

    def step(self, ids):
        self.frame += 1
        f = self.frame
        new = set()
        nx, ny, nz = self.nd
        for i in ids:
            x, y, z = ((i >> 20) & 1023) - 1, ((i >> 10) & 1023) - 1, (i & 1023) - 1
            if not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz):
                continue
            t = (z * ny + y) * nx + x
            if t not in self.info:
                continue
            if t in self.node_slot:
                self.slot_last[self.node_slot[t]] = f
                continue
            if not self._non_const(t):
                continue
            if t not in self.pending:
                self.pending.add(t)
                new.add(t)
            p = self._parent(t)
            if p is not None and p not in self.node_slot and self._non_const(p) \
                    and p not in self.pending:
                self.pending.add(p)
                new.add(p)
        for t in sorted(new):
            self.fifo.append(t)
        keys = []
        for s in range(self.nslots):
            if self.slot_node[s] is None:
                k = 0
            else:
                age = f - self.slot_last[s]
                k = 256 if age == 0 else 256 - min(age, 255)
            if k < 256:
                keys.append((k, s))
        keys.sort()
        n = min(len(self.fifo), self.max_up, len(keys))
        chosen = sorted(s for _, s in keys[:n])
        done = []
        for s in chosen:
            t = self.fifo.popleft()
            old = self.slot_node[s]
            if old is not None:
                del self.node_slot[old]
                self.tex[old, 0] &= ~np.uint32(RB)
            self.slot_node[s] = t
            self.slot_last[s] = f
            self.node_slot[t] = s
            self.pending.discard(t)
            ox, oy, oz = self._slot_origin(s)
            self.tex[t, 0] |= np.uint32(RB)
            self.tex[t, 1] = s if self.wide else ((ox + 1) << 20) | ((oy + 1) << 10) | (oz + 1)
            done.append((t, s))
        self.admitted.append(done)
        return n

    def texture(self):
        nx, ny, nz = self.nd
        return self.tex.reshape(nz, ny, nx, 2)


class RefSharedServer:
    """The same policy over a brick pool shared by several trees (one LRU
    for every tree, as the reference's BrickPool singleton serves every tree
    its pager has loaded, GigaVoxelsBrickPool.cpp 406-449, 561-682;
    GigaVoxelsDatabasePager.cpp 300-310).  A node is named by its global
    texel g = tree slot * cap + texel.

    attach  the tree takes the lowest free slot; texels that already hold a
            brick are adopted; a NON-CONST root without one becomes pending
            and joins the FIFO with the next frame's new pending nodes;
    detach  the tree's slots become empty (key 0, so they are taken first,
            as notifyDeleted moves them to the front of the LRU list) and its
            nodes lose their resident bits; its FIFO entries leave the queue,
            the others keeping their order;
    step    ids of several request lists (each with its tree slot) are
            touched, then admission runs once over all trees.
    """

    def __init__(self, slots, B, max_trees, cap, max_up, wide=False):
        self.wide = wide          # a wide pool: slots (1, 1, n), word 1 the slot
        self.slots = tuple(slots)
        self.e = tuple(b + 2 for b in B)
        self.nslots = slots[0] * slots[1] * slots[2]
        self.max_up = max_up if max_up else self.nslots
        self.max_trees, self.cap = max_trees, cap
        self.trees = [None] * max_trees
        self.node_slot = {}
        self.slot_node = [None] * self.nslots
        self.slot_last = [0] * self.nslots
        self.fifo = deque()
        self.pending = set()
        self.new = set()
        self.frame = 0
        self.admitted = []
        self.evictions = 0

    def attach(self, tex, types, bidx, texel):
        i = self.trees.index(None)
        t = dict(nd=(tex.shape[2], tex.shape[1], tex.shape[0]), tex=tex.reshape(-1, 2).copy(),
                 types=types, bidx=bidx, texel=texel, info={})
        assert t["tex"].shape[0] <= self.cap
        for d, tx in enumerate(texel):
            for n in np.nonzero(tx != NONE)[0]:
                t["info"][int(tx[n])] = (d, int(n))
        self.trees[i] = t
        base = i * self.cap
        for x in np.nonzero(t["tex"][:, 0] & RB)[0]:
            w1 = int(t["tex"][x, 1])
            sx = (((w1 >> 20) & 1023) - 1) // self.e[0]
            sy = (((w1 >> 10) & 1023) - 1) // self.e[1]
            sz = ((w1 & 1023) - 1) // self.e[2]
            s = w1 if self.wide else (sz * self.slots[1] + sy) * self.slots[0] + sx
            assert self.slot_node[s] is None, "slot taken"
            self.node_slot[base + int(x)] = s
            self.slot_node[s] = base + int(x)
        r = int(texel[0][0])
        if int(types[0][0]) == 1 and base + r not in self.node_slot:
            self.pending.add(base + r)
            self.new.add(base + r)
        return i

    def detach(self, i):
        base, t = i * self.cap, self.trees[i]
        for s in range(self.nslots):
            g = self.slot_node[s]
            if g is not None and base <= g < base + self.cap:
                t["tex"][g - base, 0] &= ~np.uint32(RB)
                del self.node_slot[g]
                self.slot_node[s] = None
                self.slot_last[s] = 0
        self.fifo = deque(g for g in self.fifo if not base <= g < base + self.cap)
        self.pending = {g for g in self.pending if not base <= g < base + self.cap}
        self.new = {g for g in self.new if not base <= g < base + self.cap}
        self.trees[i] = None

    def _split(self, g):
        i = g // self.cap
        return i, self.trees[i], g - i * self.cap

    def _non_const(self, g):
        _, t, x = self._split(g)
        d, n = t["info"][x]
        return int(t["types"][d][n]) == 1

    def _parent(self, g):
        i, t, x = self._split(g)
        d, n = t["info"][x]
        if d == 0:
            return None
        R = 1 << d
        px, py, pz = (n % R) >> 1, ((n // R) % R) >> 1, (n // (R * R)) >> 1
        v = int(t["texel"][d - 1][(pz * (R >> 1) + py) * (R >> 1) + px])
        return None if v == NONE else i * self.cap + v

    def step(self, lists):
        """lists: [(tree slot, request ids)]; returns the admitted count."""
        self.frame += 1
        f = self.frame
        new = self.new
        self.new = set()
        for i, ids in lists:
            t = self.trees[i]
            if t is None:
                continue
            nx, ny, nz = t["nd"]
            base = i * self.cap
            for q in ids:
                x, y, z = ((q >> 20) & 1023) - 1, ((q >> 10) & 1023) - 1, (q & 1023) - 1
                if not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz):
                    continue
                lx = (z * ny + y) * nx + x
                if lx not in t["info"]:
                    continue
                g = base + lx
                if g in self.node_slot:
                    self.slot_last[self.node_slot[g]] = f
                    continue
                if not self._non_const(g):
                    continue
                if g not in self.pending:
                    self.pending.add(g)
                    new.add(g)
                p = self._parent(g)
                if p is not None and p not in self.node_slot and self._non_const(p) \
                        and p not in self.pending:
                    self.pending.add(p)
                    new.add(p)
        for g in sorted(new):
            self.fifo.append(g)
        keys = []
        for s in range(self.nslots):
            if self.slot_node[s] is None:
                k = 0
            else:
                age = f - self.slot_last[s]
                k = 256 if age == 0 else 256 - min(age, 255)
            if k < 256:
                keys.append((k, s))
        keys.sort()
        n = min(len(self.fifo), self.max_up, len(keys))
        chosen = sorted(s for _, s in keys[:n])
        done = []
        for s in chosen:
            g = self.fifo.popleft()
            old = self.slot_node[s]
            if old is not None:
                del self.node_slot[old]
                _, to, xo = self._split(old)
                to["tex"][xo, 0] &= ~np.uint32(RB)
                self.evictions += 1
            self.slot_node[s] = g
            self.slot_last[s] = f
            self.node_slot[g] = s
            self.pending.discard(g)
            _, t, x = self._split(g)
            sx, sy = self.slots[0], self.slots[1]
            ox, oy, oz = (s % sx) * self.e[0], ((s // sx) % sy) * self.e[1], \
                (s // (sx * sy)) * self.e[2]
            t["tex"][x, 0] |= np.uint32(RB)
            t["tex"][x, 1] = s if self.wide else ((ox + 1) << 20) | ((oy + 1) << 10) | (oz + 1)
            done.append((g, s))
        self.admitted.append(done)
        return n

    def texture(self, i):
        t = self.trees[i]
        nx, ny, nz = t["nd"]
        return t["tex"].reshape(nz, ny, nx, 2)
