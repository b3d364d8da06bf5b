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
    def __init__(self, tex, types, bidx, texel, slots, B, max_up):
        """tex: (z, y, x, 2) node texture after the octree is built; types /
        bidx / texel: per-depth arrays of the tree; slots: pool slots per axis."""
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
            s = (z * slots[1] + y) * slots[0] + x
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
            self.tex[t, 1] = ((ox + 1) << 20) | ((oy + 1) << 10) | (oz + 1)
            done.append((t, s))
        self.admitted.append(done)
        return n

    def texture(self):
        nx, ny, nz = self.nd
        return self.tex.reshape(nz, ny, nx, 2)
