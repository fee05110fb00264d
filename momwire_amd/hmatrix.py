"""Block-low-rank (H-matrix / ACA) port solve on the GPU.

``HMatrixSolver`` is ``BSplineSolver`` with the dense fill and LU replaced by
a compressed operator and a Krylov solve, for the reference's
``HMatrixSolver`` (reference src/momwire/hmatrix.py:275-1918, _aca.py):

- the block partition of Z (momwire_amd._hpart): basis boxes, the cluster
  tree, admissible (far) and near leaf blocks -- the reference's rules and
  parameters (aca_eta, aca_leaf_size);
- near blocks dense, filled on the GPU with the dense fill's entries
  (mw_hm_fill_dense, same-edge pairs from the analytic table);
- far blocks compressed by partial-pivoted ACA to aca_tol, every block at
  once, one workgroup each (mw_hm_aca); a block whose factors would not
  compress is stored dense, as in the reference (hmatrix.py:1054-1058);
- the port columns (and, with junctions, the KCL constraint columns) solved
  by left-preconditioned restarted block GMRES to solve_tol, all columns in
  one block Krylov space (the reference's _block_gmres, hmatrix.py:177-273),
  the operator product on the GPU (mw_hm_apply);
- the KCL Schur step and the readout of the dense path.

The preconditioner is block Jacobi over the cluster tree's leaves (their
diagonal blocks inverted on the GPU, mw_hm_invert) where the reference
factors the whole near field with a sparse LU (hmatrix.py:95-112, 1468-1493):
the answer is the same to solve_tol; only the iteration count can differ.

Free space only: a ground plane or wire loading raises NotImplementedError.
"""

import ctypes
import math

import numpy as np

from . import _lib
from ._hpart import Partition
from ._port_solution import PortSolution
from .bspline import BSplineSolver, _to_host, _torch

HBLOCK = np.dtype([("m", np.int32), ("n", np.int32), ("rank", np.int32), ("cap", np.int32),
                   ("row_off", np.int64), ("col_off", np.int64), ("data_off", np.int64),
                   ("aux_off", np.int64), ("t_off", np.int64), ("row_start", np.int64),
                   ("col_start", np.int64)])
TITEM = np.dtype([("row0", np.int32), ("nrows", np.int32), ("k", np.int32), ("pad", np.int32),
                  ("col_off", np.int64)])
RITEM = np.dtype([("row0", np.int64), ("nrows", np.int32), ("ent0", np.int32), ("nent", np.int32),
                  ("pad", np.int32)])
assert HBLOCK.itemsize == 72 and TITEM.itemsize == 24 and RITEM.itemsize == 24


def _r_items(part, keep_block):
    """Phase-2 work of the product: per row chunk, its (block, row offset)
    entries in block order, blocks with keep_block False left out."""
    sel = keep_block[part.ent_block]
    chunk = part.ent_chunk[sel]
    order = np.argsort(chunk, kind="stable")
    ent = np.stack([part.ent_block[sel][order], part.ent_ro[sel][order]], axis=1)
    counts = np.bincount(chunk, minlength=len(part.chunk_start))
    items = np.zeros(len(part.chunk_start), dtype=RITEM)
    items["row0"] = part.chunk_start
    items["nrows"] = part.chunk_rows
    items["ent0"] = np.concatenate([[0], np.cumsum(counts)[:-1]])
    items["nent"] = counts
    return items, np.ascontiguousarray(ent.astype(np.int32))


def _t_items(part, blocks, lr_ids, n):
    """Phase-1 work of the product: the compressed blocks grouped by column
    cluster, their (block, rank) rows listed group after group, and one item
    per 64 rows of a group.  Returns (items TITEM, rows int32 [R][2])."""
    if not len(lr_ids):
        return np.zeros(0, dtype=TITEM), np.zeros((0, 2), dtype=np.int32)
    keyc = part.col_start[lr_ids] * (n + 1) + blocks["n"][lr_ids]
    order = np.argsort(keyc, kind="stable")
    grp = lr_ids[order]
    rk = blocks["rank"][grp].astype(np.int64)
    # rows: every block's ranks 0 .. r-1, blocks in group order
    first = np.concatenate([[0], np.cumsum(rk)[:-1]])
    rows = np.stack([np.repeat(grp, rk), np.arange(rk.sum()) - np.repeat(first, rk)], axis=1)
    # groups: runs of equal key; their row spans and their 64-row tiles
    newg = np.concatenate([[True], np.diff(keyc[order]) != 0])
    gstart_blk = np.flatnonzero(newg)
    g_rows = np.add.reduceat(rk, gstart_blk)
    g_base = np.concatenate([[0], np.cumsum(g_rows)[:-1]])
    ntile = -(-g_rows // 64)
    tg = np.repeat(np.arange(len(g_rows)), ntile)
    t_in = np.arange(ntile.sum()) - np.repeat(np.concatenate([[0], np.cumsum(ntile)[:-1]]), ntile)
    items = np.zeros(len(tg), dtype=TITEM)
    items["row0"] = g_base[tg] + 64 * t_in
    items["nrows"] = np.minimum(64, g_rows[tg] - 64 * t_in)
    lead = grp[gstart_blk]
    items["k"] = blocks["n"][lead][tg]
    items["col_off"] = blocks["col_off"][lead][tg]
    return items, np.ascontiguousarray(rows.astype(np.int32))


class HOperator:
    """The compressed operator on the device: block table, index pools and
    store, and the work lists of its product (and of the preconditioner).
    The product runs in leaf order (the partition's perm)."""

    def __init__(self, solver, part, tol, k):
        torch = _torch()
        self.torch = torch
        lib = _lib.load()
        self.lib = lib
        dev = torch.device("cuda")
        D = solver._device_tables()
        st = torch.cuda.current_stream().cuda_stream
        n = part.n
        self.n = n
        self.part = part
        blocks = np.zeros(len(part.blocks), dtype=HBLOCK)
        blocks["m"] = part.blocks[:, 1]
        blocks["n"] = part.blocks[:, 3]
        blocks["row_off"] = part.blocks[:, 0]
        blocks["col_off"] = part.blocks[:, 2]
        blocks["row_start"] = part.row_start
        blocks["col_start"] = part.col_start
        far = part.blocks[:, 4] == 1
        m64 = blocks["m"].astype(np.int64)
        n64 = blocks["n"].astype(np.int64)
        # far blocks: factor capacity up to the rank at which dense storage
        # is as cheap (the block then goes dense anyway), at most rank_cap
        r_dense = -(-(m64 * n64) // (m64 + n64))
        cap = np.where(far, np.minimum(r_dense, solver.aca_rank_cap), 0).astype(np.int32)
        blocks["cap"] = cap
        blocks["rank"] = np.where(far, 0, -1)
        size = np.where(far, cap.astype(np.int64) * (m64 + n64), m64 * n64)
        off = np.zeros(len(blocks), dtype=np.int64)
        np.cumsum(size[:-1], out=off[1:])
        blocks["data_off"] = off
        flag_sz = np.where(far, m64 + n64, 0)
        aux = np.zeros(len(blocks), dtype=np.int64)
        np.cumsum(flag_sz[:-1], out=aux[1:])
        blocks["aux_off"] = aux

        cache = D.keep.setdefault("hm_pools", {})
        key = id(part)
        if key not in cache:
            up0 = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
            cache.clear()
            cache[key] = dict(rows=up0(part.rows_pool), cols=up0(part.cols_pool),
                              lrows=up0(part.rows_lpool), lcols=up0(part.cols_lpool),
                              sigma=up0(part.sigma), colpos=up0(part.colpos),
                              perm=up0(part.perm), pos=up0(part.pos))
        pools = cache[key]
        self.pools = pools
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.wing_seg = D.keep["hm_wing_seg"]
        self.wing_slot = D.keep["hm_wing_slot"]
        store = torch.empty(int(size.sum()), dtype=torch.complex128, device=dev)
        d_blocks = up(blocks.view(np.uint8))
        args = (D.mw, self.wing_seg.data_ptr(), self.wing_slot.data_ptr(), k, k * solver.c,
                solver.eps, solver.mu)

        # near blocks: dense, same-edge pairs from the analytic table
        _lib.check(lib.mw_same_edge_table(D.mw, k, st), "same-edge table")
        self._fill_dense(args, d_blocks, blocks, np.flatnonzero(~far), 1, store)
        # far blocks: ACA, all at once, the largest first (they finish last)
        far_ids = np.flatnonzero(far).astype(np.int32)
        far_ids = far_ids[np.argsort(-(m64 + n64)[far_ids], kind="stable")]
        if far_ids.size:
            flags = torch.empty(int(flag_sz.sum()), dtype=torch.uint8, device=dev)
            fl = up(far_ids)
            # the large blocks (512-thread workgroups) on a second stream, so
            # that they run beside the rest instead of ahead of them
            n_large = int(((m64 + n64)[far_ids] >= 1024).sum())
            aca = lambda lst, cnt, large, stream: _lib.check(lib.mw_hm_aca(  # noqa: E731
                *args, pools["rows"].data_ptr(), pools["cols"].data_ptr(), d_blocks.data_ptr(),
                lst, cnt, large, float(tol), store.data_ptr(), flags.data_ptr(), stream), "ACA")
            if n_large:
                cur = torch.cuda.current_stream()
                side = torch.cuda.Stream()
                side.wait_stream(cur)
                aca(fl.data_ptr(), n_large, n_large, side.cuda_stream)
            if far_ids.size > n_large:
                aca(fl.data_ptr() + 4 * n_large, int(far_ids.size) - n_large, 0, st)
            if n_large:
                cur.wait_stream(side)
            got, = _to_host(d_blocks)
            blocks["rank"] = got.view(HBLOCK)["rank"]
            del flags
        # far blocks that did not compress: dense, off-edge kernel
        dense_far = np.flatnonzero(far & (blocks["rank"] < 0))
        if dense_far.size:
            extra = (m64 * n64)[dense_far]
            base = store.numel()
            blocks["data_off"][dense_far] = base + np.concatenate([[0], np.cumsum(extra)[:-1]])
            blocks["cap"][dense_far] = 0
            store = torch.cat([store, torch.empty(int(extra.sum()), dtype=torch.complex128,
                                                  device=dev)])
            d_blocks = up(blocks.view(np.uint8))
            self._fill_dense(args, d_blocks, blocks, dense_far, 0, store)
        lr = far & (blocks["rank"] > 0)
        lr_ids = np.flatnonzero(lr)
        t_off = np.zeros(len(blocks), dtype=np.int64)
        if lr_ids.size:
            t_off[lr_ids] = np.concatenate([[0], np.cumsum(blocks["rank"][lr_ids])[:-1]])
        blocks["t_off"] = t_off
        self.t_rows = int(blocks["rank"][lr_ids].sum()) if lr_ids.size else 0
        # phase 1 of the product: the V rows of blocks sharing a column
        # cluster, 64 at a time
        ti, tr = _t_items(part, blocks, lr_ids, n)
        self.t_items = up(ti.view(np.uint8)) if len(ti) else None
        self.n_t = len(ti)
        self.t_rowlist = up(tr) if len(ti) else None
        # phase 2: every row chunk over the blocks that hold it (compressed
        # blocks of rank 0 add nothing)
        keep = ~(far & (blocks["rank"] == 0))
        r_items, r_ent = _r_items(part, keep)
        self.r_items, self.r_ent = up(r_items.view(np.uint8)), up(r_ent)
        self.n_r = len(r_items)
        self.blocks = blocks
        self.d_blocks = up(blocks.view(np.uint8))
        self.store = store
        self._precond_built = False

    def _fill_dense(self, args, d_blocks, blocks, ids, table, store):
        """Dense blocks in leaf order (a leaf's members keep ascending order,
        so near blocks read the same either way)."""
        if not len(ids):
            return
        ids = np.asarray(ids, dtype=np.int64)
        nit = -(-(blocks["m"][ids].astype(np.int64) * blocks["n"][ids]) // 256)
        first = np.arange(nit.sum()) - np.repeat(np.concatenate([[0], np.cumsum(nit)[:-1]]), nit)
        items = (np.repeat(ids, nit) << 32) | (256 * first)
        d_items = self.torch.from_numpy(items).to(store.device)
        _lib.check(self.lib.mw_hm_fill_dense(*args, self.pools["lrows"].data_ptr(),
                                             self.pools["lcols"].data_ptr(), d_blocks.data_ptr(),
                                             d_items.data_ptr(), int(items.size), table,
                                             store.data_ptr(), self.torch.cuda.current_stream()
                                             .cuda_stream), "dense blocks")

    # ------------------------------------------------------------------
    def stats(self):
        """The reference's HMatrix.stats() (_aca.py:255-270); blocks stored
        dense after ACA count as near."""
        b = self.blocks
        lr = b["rank"] >= 0
        dense = ~lr
        near_area = int((b["m"][dense].astype(np.int64) * b["n"][dense]).sum())
        ranks = b["rank"][lr]
        storage = near_area + int((ranks.astype(np.int64) * (b["m"][lr] + b["n"][lr])).sum())
        return {
            "n": self.n, "n_near": int(dense.sum()), "n_far": int(lr.sum()),
            "storage": storage, "dense_storage": self.n * self.n,
            "compression": storage / (self.n * self.n), "near_area": near_area,
            "max_rank": int(ranks.max()) if ranks.size else 0,
            "mean_rank": float(ranks.mean()) if ranks.size else 0.0,
        }

    def to_dense(self):
        """The operator as a dense numpy matrix in basis order (validation,
        small n): dense blocks as stored, compressed blocks as U V
        (_aca.py:280-287)."""
        store, = _to_host(self.store)
        part = self.part
        Z = np.zeros((self.n, self.n), dtype=np.complex128)
        for b, B in enumerate(self.blocks):
            m, n, o = int(B["m"]), int(B["n"]), int(B["data_off"])
            ro, co = int(B["row_off"]), int(B["col_off"])
            if B["rank"] < 0:
                I, J = part.rows_lpool[ro:ro + m], part.cols_lpool[co:co + n]
                Z[np.ix_(I, J)] = store[o:o + m * n].reshape(m, n)
            else:
                I, J = part.rows_pool[ro:ro + m], part.cols_pool[co:co + n]
                r, cap = int(B["rank"]), int(B["cap"])
                Ut = store[o:o + cap * m].reshape(cap, m)[:r]
                V = store[o + cap * m:o + cap * (m + n)].reshape(cap, n)[:r]
                Z[np.ix_(I, J)] = Ut.T @ V
        return Z

    def apply_leaf(self, X):
        """H X for a (n, s) complex128 device tensor in leaf order."""
        torch = self.torch
        X = X.contiguous()
        s = X.shape[1]
        Y = torch.empty_like(X)
        T = (torch.empty((max(self.t_rows, 1), s), dtype=torch.complex128, device=X.device)
             if self.n_t else None)
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.mw_hm_apply(
            self.d_blocks.data_ptr(), self.store.data_ptr(),
            self.t_items.data_ptr() if self.n_t else None, self.n_t,
            self.t_rowlist.data_ptr() if self.n_t else None, self.pools["colpos"].data_ptr(),
            self.r_items.data_ptr(), self.n_r, self.r_ent.data_ptr(),
            self.pools["sigma"].data_ptr(), X.data_ptr(), s, s,
            T.data_ptr() if T is not None else None, Y.data_ptr(), s, st), "H X")
        return Y

    def apply(self, X):
        """H X for a (n, s) complex128 device tensor in basis order."""
        perm, pos = self.pools["perm"], self.pools["pos"]
        return self.apply_leaf(X[perm])[pos]

    def _build_precond(self):
        """Block Jacobi: the inverses of the leaf diagonal blocks, in leaf
        order."""
        torch = self.torch
        part, blocks = self.part, self.blocks
        dev = self.store.device
        diag = np.flatnonzero((part.blocks[:, 4] == 0) & (part.row_start == part.col_start)
                              & (blocks["m"] == blocks["n"]))
        diag = diag[np.argsort(part.row_start[diag], kind="stable")]  # leaf order
        pb = np.zeros(len(diag), dtype=HBLOCK)
        pb["m"] = blocks["m"][diag]
        pb["n"] = blocks["n"][diag]
        pb["rank"] = -1
        pb["row_start"] = part.row_start[diag]
        pb["col_start"] = part.col_start[diag]
        sz = pb["m"].astype(np.int64) * pb["n"]
        pb["data_off"] = np.concatenate([[0], np.cumsum(sz)[:-1]])
        pb["aux_off"] = np.concatenate([[0], np.cumsum(2 * sz)[:-1]])
        ends = pb["row_start"] + pb["m"]
        if (pb["row_start"][0] != 0 or ends[-1] != self.n
                or not np.array_equal(pb["row_start"][1:], ends[:-1])):
            raise RuntimeError("the leaf diagonal blocks do not cover every basis once")
        src = (np.repeat(blocks["data_off"][diag] - pb["data_off"], sz)
               + np.arange(int(sz.sum())))
        pstore = self.store[torch.from_numpy(src).to(dev)].contiguous()
        work = torch.empty(int(2 * sz.sum()), dtype=torch.complex128, device=dev)
        info = torch.zeros(len(diag), dtype=torch.int32, device=dev)
        d_pb = torch.from_numpy(pb.view(np.uint8).copy()).to(dev)
        lst = torch.arange(len(diag), dtype=torch.int32, device=dev)
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.mw_hm_invert(d_pb.data_ptr(), lst.data_ptr(), len(diag),
                                         pstore.data_ptr(), work.data_ptr(), info.data_ptr(), st),
                   "block-Jacobi inverses")
        bad, = _to_host(info)
        if bad.any():
            raise np.linalg.LinAlgError("a leaf diagonal block of Z is singular")
        # one entry per row chunk: the inverse of its leaf
        leaf_of = np.searchsorted(pb["row_start"], part.chunk_start, side="right") - 1
        items = np.zeros(len(part.chunk_start), dtype=RITEM)
        items["row0"] = part.chunk_start
        items["nrows"] = part.chunk_rows
        items["ent0"] = np.arange(len(part.chunk_start))
        items["nent"] = 1
        ent = np.stack([leaf_of, part.chunk_start - pb["row_start"][leaf_of]], axis=1)
        self.p_blocks, self.p_store = d_pb, pstore
        self.p_items = torch.from_numpy(items.view(np.uint8).copy()).to(dev)
        self.p_ent = torch.from_numpy(np.ascontiguousarray(ent.astype(np.int32))).to(dev)
        self._precond_built = True

    def precond_leaf(self, R):
        """M^{-1} R in leaf order, M = the leaf diagonal blocks of Z."""
        torch = self.torch
        if not self._precond_built:
            self._build_precond()
        R = R.contiguous()
        s = R.shape[1]
        Y = torch.empty_like(R)
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.mw_hm_apply(
            self.p_blocks.data_ptr(), self.p_store.data_ptr(), None, 0, None, None,
            self.p_items.data_ptr(), len(self.part.chunk_start), self.p_ent.data_ptr(), None,
            R.data_ptr(), s, s, None, Y.data_ptr(), s, st), "M^-1 R")
        return Y

    def precond(self, R):
        """M^{-1} R in basis order."""
        perm, pos = self.pools["perm"], self.pools["pos"]
        return self.precond_leaf(R[perm])[pos]

    def solve(self, B, rtol):
        """Block GMRES for H X = B (basis order), run in leaf order."""
        perm, pos = self.pools["perm"], self.pools["pos"]
        X, iters = block_gmres(self.apply_leaf, self.precond_leaf, B[perm].contiguous(), rtol)
        return X[pos].contiguous(), iters


def _qr_tall(W):
    """W = Q R for a tall block (N x s): CholeskyQR2 -- two passes of the
    Gram matrix's Cholesky factor, GEMMs on the MFMA units instead of a
    column-by-column Householder sweep.  A Gram matrix that is not
    numerically positive definite (nearly dependent columns) falls back to
    Householder QR."""
    import torch

    Q, R = W, None
    for _ in range(2):
        L, info = torch.linalg.cholesky_ex(Q.conj().T @ Q)
        if int(info) != 0:
            return torch.linalg.qr(W)
        Rk = L.conj().T
        Q = torch.linalg.solve_triangular(Rk, Q, upper=True, left=False)
        R = Rk if R is None else Rk @ R
    return Q, R


def block_gmres(op, prec, B, rtol, restart=None, maxiter=2000):
    """Left-preconditioned restarted block GMRES for op(X) = B, every column
    in one block Krylov space (the reference's _block_gmres,
    hmatrix.py:177-273): preconditioned initial guess M^{-1} B, CGS2
    orthogonalisation, stopping on the per-column preconditioned residual
    relative to |M^{-1} b_j|.  The small least-squares problem over the block
    Hessenberg is solved afresh at each step from a thin QR of the Hessenberg
    (the reference updates one incrementally with block rotations; the
    solution and the residual test are the same), so every factorisation is
    a CholeskyQR of a tall block.  Device tensors throughout.  Returns (X,
    iterations)."""
    import torch

    N, s = B.shape
    if restart is None:
        restart = 50
    Bt = prec(B)
    bnorms = torch.linalg.vector_norm(Bt, dim=0)
    bnorms = torch.where(bnorms == 0, torch.ones_like(bnorms), bnorms)
    X = Bt.clone()
    m = min(restart, N)
    total = 0
    dt, dev = B.dtype, B.device
    for _ in range(maxiter // m + 1):
        R = Bt - prec(op(X))
        if bool((torch.linalg.vector_norm(R, dim=0) <= rtol * bnorms).all()):
            break
        Q, beta = _qr_tall(R)
        V = torch.empty((N, (m + 1) * s), dtype=dt, device=dev)
        V[:, :s] = Q
        Hs = torch.zeros(((m + 1) * s, m * s), dtype=dt, device=dev)
        E = torch.zeros(((m + 1) * s, s), dtype=dt, device=dev)
        E[:s] = beta
        Yk = None
        psz = 0
        for k in range(m):
            total += 1
            p = (k + 1) * s
            W = prec(op(V[:, p - s:p]))
            Vp = V[:, :p]
            H1 = Vp.conj().T @ W
            W = W - Vp @ H1
            H2 = Vp.conj().T @ W
            W = W - Vp @ H2
            Qk, Hkk = _qr_tall(W)
            V[:, p:p + s] = Qk
            Hs[:p, p - s:p] = H1 + H2
            Hs[p:p + s, p - s:p] = Hkk
            Hk = Hs[:p + s, :p]
            Qh, Rh = _qr_tall(Hk)
            Ek = E[:p + s]
            diag = torch.abs(torch.diagonal(Rh))
            if bool((diag > 1e-14 * max(float(diag.max()), 1.0)).all()):
                Yk = torch.linalg.solve_triangular(Rh, Qh.conj().T @ Ek, upper=True)
            else:  # near-breakdown: least squares on the host
                Yk = torch.from_numpy(np.linalg.lstsq(Hk.cpu().numpy(), Ek.cpu().numpy(),
                                                      rcond=None)[0]).to(dev)
            psz = p
            rk = torch.linalg.vector_norm(Ek - Hk @ Yk, dim=0)
            if bool((rk <= rtol * bnorms).all()):
                break
        X = X + V[:, :psz] @ Yk
        if bool((rk <= rtol * bnorms).all()):
            break
    return X, total


class HMatrixSolver(BSplineSolver):
    """BSplineSolver whose port solve runs on the block-low-rank operator
    (the reference's HMatrixSolver: same constructor and parameters,
    hmatrix.py:1872-1918; aca_rank_cap bounds a far block's ACA rank -- a
    block that reaches it is stored dense, exact)."""

    def __init__(self, *args, aca_eta=1.0, aca_leaf_size=32, aca_tol=1e-4, solve_tol=1e-6,
                 hmatrix_use_accel=True, precond_eta=None, swept_dense_max_bases=None,
                 aca_rank_cap=128, **kwargs):
        super().__init__(*args, **kwargs)
        self.aca_eta = float(aca_eta)
        self.aca_leaf_size = int(aca_leaf_size)
        self.aca_tol = float(aca_tol)
        self.solve_tol = float(solve_tol)
        self.hmatrix_use_accel = bool(hmatrix_use_accel)
        self.precond_eta = 0.5 * self.aca_eta if precond_eta is None else float(precond_eta)
        self.swept_dense_max_bases = swept_dense_max_bases
        self.aca_rank_cap = int(aca_rank_cap)
        if self.aca_leaf_size < 1:
            raise ValueError("aca_leaf_size must be positive")
        if not self.aca_tol > 0 or not self.solve_tol > 0:
            raise ValueError("aca_tol and solve_tol must be positive")
        if self.aca_rank_cap < 1:
            raise ValueError("aca_rank_cap must be positive")
        self._hm_partition = None
        self._hmatrix = None
        self._last_solve_iters = None

    def _hm_unsupported(self):
        if self.ground_z is not None:
            raise NotImplementedError("the block-low-rank operator is built for free space "
                                      "(solve a grounded structure with BSplineSolver)")
        if self._loading_active:
            raise NotImplementedError("wire loading with the block-low-rank operator "
                                      "(solve it with BSplineSolver)")

    def build_partition(self, eta=None, leaf_size=None):
        """The block partition (memoised per eta and leaf size): a Partition
        with the leaf blocks, their cluster members and counts."""
        eta = self.aca_eta if eta is None else float(eta)
        leaf = self.aca_leaf_size if leaf_size is None else int(leaf_size)
        c = self._hm_partition
        if c is not None and c[0] == eta and c[1] == leaf:
            return c[2]
        part = Partition(self._build(), eta, leaf)
        self._hm_partition = (eta, leaf, part)
        return part

    def _device_tables(self):
        D = self._device_base()
        if "hm_wing_seg" not in D.keep:
            torch = D.torch
            mesh = self._mesh
            D.keep["hm_wing_seg"] = torch.from_numpy(
                np.ascontiguousarray(mesh.supp, dtype=np.int32)).to(D.dev)
            D.keep["hm_wing_slot"] = torch.from_numpy(
                np.ascontiguousarray(mesh.bw_slot, dtype=np.int32)).to(D.dev)
        return D

    def build_hmatrix(self, eta=None, leaf_size=None, tol=None, k=None):
        """The compressed operator at wavenumber k (default: the solver's)."""
        self._hm_unsupported()
        part = self.build_partition(eta, leaf_size)
        return HOperator(self, part, self.aca_tol if tol is None else float(tol),
                         self.k if k is None else float(k))

    def _solve(self, k=None, timers=None):
        """Port solve on the compressed operator; returns device (Y, W)."""
        torch = _torch()
        lib = _lib.load()
        D = self._device_tables()
        D.post_buffers()
        H = self.build_hmatrix(k=k)
        self._hmatrix = H
        st = torch.cuda.current_stream().cuda_stream
        nb, ncols = D.nb, D.ncols
        B = torch.empty((nb, max(ncols, 1)), dtype=torch.complex128, device=D.dev)
        _lib.check(lib.mw_zscatter_cols(nb, ncols, D.col_max, D.col_idx.data_ptr(),
                                        D.col_val.data_ptr(), B.data_ptr(), B.stride(0), st),
                   "right-hand sides")
        X, iters = H.solve(B[:, :ncols], self.solve_tol)
        self._last_solve_iters = [iters] * D.n_ports
        X = X.contiguous()
        D.info.zero_()
        self._post(D, X.data_ptr(), X.stride(0), D.Y.data_ptr(), st)
        return D.Y[:D.n_ports, :D.n_ports], X[:, :D.n_ports]

    def compute_port_solution(self, same_edge_prep=None) -> PortSolution:
        """Every port column from one operator build and one block-GMRES
        solve (reference hmatrix.py:1552-1605); columns converged to
        solve_tol."""
        Y, W = self._solve()
        y, coeffs = _to_host(Y, W)
        info, = _to_host(self._dev.info)
        if info[1] != 0:
            raise np.linalg.LinAlgError("KCL Schur complement is singular")
        return PortSolution(y=y, coeffs=coeffs, port_currents=y,
                            basis=("bspline-gpu", self._mesh, id(self)))

    def compute_y_matrix(self):
        return self.compute_port_solution().y

    def solve_graph(self, warm=True):
        raise NotImplementedError("the block-low-rank solve is not captured as a graph")
