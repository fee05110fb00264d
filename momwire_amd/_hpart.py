"""Block partition of the impedance matrix for the block-low-rank operator.

Host-side geometry only: per-basis bounding boxes, the binary cluster tree
over them, and the admissible (far, compressible) / inadmissible (near,
dense) leaf blocks of Z.  The rules are the reference's, kept alike so that
its parameters keep their meaning and the same structure gets the same
partition (reference src/momwire/_aca.py:61-134, hmatrix.py:339-367,
939-980):

- a basis's box spans the end points of the segments its live wings cover;
- a cluster is bisected along its box's longest axis at the median of its
  members' box centres (an index-count split when the median leaves one side
  empty), down to ``leaf_size`` members;
- a block (s, t) is far when min(diam s, diam t) <= eta * dist(s, t) with
  dist the gap between the boxes (0 when they touch); the recursion splits
  whichever side is not a leaf and stops at two leaves (near).

A cluster's members stay in ascending basis order: the order in which the
compression walks a block's rows and columns.
"""

import numpy as np


class Cluster:
    __slots__ = ("indices", "lo", "hi", "left", "right")

    def __init__(self, indices, lo, hi):
        self.indices = indices
        self.lo = lo
        self.hi = hi
        self.left = None
        self.right = None

    @property
    def is_leaf(self):
        return self.left is None

    @property
    def diam(self):
        return float(np.linalg.norm(self.hi - self.lo))

    def leaves(self):
        stack = [self]
        while stack:
            c = stack.pop()
            if c.is_leaf:
                yield c
            else:
                stack.append(c.right)
                stack.append(c.left)


def basis_boxes(mesh):
    """(lo, hi), each (n_basis, 3): the box of every basis's live wings."""
    seg = mesh.seg
    lo = np.empty((mesh.n_basis, 3))
    hi = np.empty((mesh.n_basis, 3))
    for m in range(mesh.n_basis):
        segs = np.unique(mesh.supp[m][mesh.supp[m] >= 0])
        pts = np.vstack([seg[segs, 0:3], seg[segs, 3:6]])
        lo[m] = pts.min(axis=0)
        hi[m] = pts.max(axis=0)
    return lo, hi


def cluster_tree(box_lo, box_hi, leaf_size):
    root_idx = np.arange(box_lo.shape[0], dtype=np.int64)

    def make(idx):
        return Cluster(idx, box_lo[idx].min(axis=0), box_hi[idx].max(axis=0))

    root = make(root_idx)
    stack = [root]
    while stack:
        node = stack.pop()
        idx = node.indices
        if idx.size <= leaf_size:
            continue
        centres = 0.5 * (box_lo[idx] + box_hi[idx])
        axis = int(np.argmax(node.hi - node.lo))
        split = float(np.median(centres[:, axis]))
        left = centres[:, axis] <= split
        if left.all() or not left.any():
            order = np.argsort(centres[:, axis], kind="stable")
            left = np.zeros(idx.size, dtype=bool)
            left[order[: idx.size // 2]] = True
        node.left = make(idx[left])
        node.right = make(idx[~left])
        stack.append(node.left)
        stack.append(node.right)
    return root


def _gap(a, b):
    d = np.maximum(np.maximum(a.lo - b.hi, b.lo - a.hi), 0.0)
    return float(np.linalg.norm(d))


def is_far(a, b, eta):
    d = _gap(a, b)
    return d != 0.0 and min(a.diam, b.diam) <= eta * d


def block_tree(root, eta):
    """(far, near): lists of (Cluster, Cluster) leaf blocks."""
    far, near = [], []
    stack = [(root, root)]
    while stack:
        s, t = stack.pop()
        if is_far(s, t, eta):
            far.append((s, t))
        elif s.is_leaf and t.is_leaf:
            near.append((s, t))
        elif s.is_leaf:
            stack += [(s, t.left), (s, t.right)]
        elif t.is_leaf:
            stack += [(s.left, t), (s.right, t)]
        else:
            stack += [(s.left, t.left), (s.left, t.right), (s.right, t.left), (s.right, t.right)]
    return far, near


class Partition:
    """The leaf blocks in a canonical order (by first row, first column),
    with the members of every block's row and column clusters laid out in
    two index pools.

    blocks: (nb, 5) int64 -- row offset into rows_pool, rows, column offset
    into cols_pool, columns, far flag."""

    def __init__(self, mesh, eta, leaf_size):
        lo, hi = basis_boxes(mesh)
        self.root = cluster_tree(lo, hi, int(leaf_size))
        far, near = block_tree(self.root, float(eta))
        rows = [(s.indices, t.indices, 1) for s, t in far]
        rows += [(s.indices, t.indices, 0) for s, t in near]
        rows.sort(key=lambda r: (int(r[0][0]), int(r[1][0]), r[0].size, r[1].size))
        desc, rp, cp = [], [], []
        nr = nc = 0
        for I, J, f in rows:
            desc.append((nr, I.size, nc, J.size, f))
            rp.append(I)
            cp.append(J)
            nr += I.size
            nc += J.size
        self.n = mesh.n_basis
        self.blocks = np.array(desc, dtype=np.int64).reshape(-1, 5)
        self.rows_pool = np.concatenate(rp).astype(np.int32)
        self.cols_pool = np.concatenate(cp).astype(np.int32)
        self.leaves = [c.indices for c in self.root.leaves()]
        self.n_far = int(self.blocks[:, 4].sum())
        self.n_near = len(self.blocks) - self.n_far
        # leaf order: the leaves' members one after another; every cluster is
        # then one contiguous range of it (a leaf's members keep their
        # ascending order, so a leaf block reads the same in both orders)
        self.perm = np.concatenate(self.leaves).astype(np.int64)
        self.pos = np.empty(self.n, dtype=np.int64)
        self.pos[self.perm] = np.arange(self.n)
        # per block: first leaf-order position of its row and column clusters
        self.row_start = np.array([self.pos[self.block_rows(b)].min() for b in
                                   range(len(self.blocks))], dtype=np.int64)
        self.col_start = np.array([self.pos[self.block_cols(b)].min() for b in
                                   range(len(self.blocks))], dtype=np.int64)
        nb = len(self.blocks)
        m = self.blocks[:, 1]
        n = self.blocks[:, 3]
        # every block's members in leaf order (dense blocks are stored so)
        self.rows_lpool = np.concatenate(
            [self.perm[self.row_start[b]:self.row_start[b] + m[b]] for b in range(nb)]
        ).astype(np.int32)
        self.cols_lpool = np.concatenate(
            [self.perm[self.col_start[b]:self.col_start[b] + n[b]] for b in range(nb)]
        ).astype(np.int32)
        # compressed blocks keep ascending member order: sigma[row_off + t] is
        # leaf-order row t's position among the block's rows, colpos the
        # leaf-order position of each column member
        self.sigma = np.concatenate(
            [np.searchsorted(self.block_rows(b), self.rows_lpool[self.blocks[b, 0]:
                                                                 self.blocks[b, 0] + m[b]])
             for b in range(nb)]).astype(np.int32)
        self.colpos = self.pos[self.cols_pool].astype(np.int32)
        # row chunks of at most 32 leaf-order rows, never across a leaf
        starts, sizes = [], []
        at = 0
        for leaf in self.leaves:
            for o in range(0, leaf.size, 32):
                starts.append(at + o)
                sizes.append(min(32, leaf.size - o))
            at += leaf.size
        self.chunk_start = np.array(starts, dtype=np.int64)
        self.chunk_rows = np.array(sizes, dtype=np.int32)
        # (chunk, block, row offset) for every block over every chunk it holds,
        # in block order
        lo = np.searchsorted(self.chunk_start, self.row_start)
        hi = np.searchsorted(self.chunk_start, self.row_start + m)
        cnt = hi - lo
        self.ent_block = np.repeat(np.arange(nb), cnt).astype(np.int32)
        first = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
        self.ent_chunk = (np.arange(cnt.sum()) + first).astype(np.int64)
        self.ent_ro = (self.chunk_start[self.ent_chunk]
                       - self.row_start[self.ent_block]).astype(np.int32)

    def block_rows(self, b):
        o, m = self.blocks[b, 0], self.blocks[b, 1]
        return self.rows_pool[o:o + m]

    def block_cols(self, b):
        o, n = self.blocks[b, 2], self.blocks[b, 3]
        return self.cols_pool[o:o + n]
