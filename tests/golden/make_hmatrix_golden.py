"""Write golden vectors for the block-low-rank (H-matrix / ACA) port solve.

Runs ONLY in the development container, where the reference checkout is
mounted read-only at /root/reference.  It imports the reference's
HMatrixSolver (pure numpy: no compiled accelerator is loaded) and records,
for each structure, what ``HMatrixSolver.compute_port_solution()`` returns
(reference src/momwire/hmatrix.py:1552-1605) with its default parameters
(aca_eta 1, aca_leaf_size 32, aca_tol 1e-4, solve_tol 1e-6,
hmatrix.py:1872-1918):

- ``y``: the port admittance matrix;
- ``coeffs``: the first coefficient columns and the column sum;
- the block partition it built (hmatrix.py:939-980): every leaf block as
  (first row id, rows, first column id, columns, far flag) with the cluster
  members listed in ``rows_pool`` / ``cols_pool``, sorted ascending as the
  reference holds them;
- ``aca_rank``: the ACA rank of every far block (-1: stored dense because
  the factors did not compress, hmatrix.py:1054-1058), in block order.

    python tests/golden/make_hmatrix_golden.py [name ...]
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, "/root/reference/src")

from momwire.hmatrix import HMatrixSolver  # noqa: E402

from momwire_amd import workloads  # noqa: E402

CASES = {
    "hm_array4x4": lambda: workloads.dipole_array(4, 4, 15, seed=2),
    "hm_array16x16": lambda: workloads.dipole_array(16, 16, 15, seed=2),
    "hm_chain120": lambda: workloads.wire_chain(120, 4, feed_every=10),
}


def _blocks(part):
    """Leaf blocks in a canonical order: sorted by (first row, first column)."""
    rows = []
    for far, lst in ((1, part["far"]), (0, part["near"])):
        for s, t in lst:
            rows.append((int(s.indices[0]), int(t.indices[0]), s.indices, t.indices, far))
    rows.sort(key=lambda r: (r[0], r[1], r[2].size, r[3].size))
    return rows


def record(name):
    kw = CASES[name]()
    s = HMatrixSolver(**kw)
    t0 = time.time()
    sol = s.compute_port_solution()
    dt = time.time() - t0
    H = s._hmatrix
    part = s.build_partition()
    blocks = _blocks(part)
    # ACA outcome per far block: rank of its factors, -1 when stored dense
    far_rank = {}
    for I, J, U, V in H.far:
        far_rank[(int(I[0]), int(J[0]), I.size, J.size)] = U.shape[1]
    desc, rpool, cpool, ranks = [], [], [], []
    for r0, c0, I, J, far in blocks:
        desc.append((len(rpool), I.size, len(cpool), J.size, far))
        rpool.extend(I.tolist())
        cpool.extend(J.tolist())
        if far:
            ranks.append(far_rank.get((r0, c0, I.size, J.size), -1))
    cols = np.arange(min(4, sol.coeffs.shape[1]))
    out = dict(
        y=sol.y, coeffs_cols=cols, coeffs=sol.coeffs[:, cols],
        coeffs_colsum=sol.coeffs.sum(axis=1),
        blocks=np.array(desc, dtype=np.int64), rows_pool=np.array(rpool, dtype=np.int32),
        cols_pool=np.array(cpool, dtype=np.int32), aca_rank=np.array(ranks, dtype=np.int32),
        n=np.int64(H.n), gmres_iters=np.int64(s._last_solve_iters[0]),
    )
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    st = H.stats()
    print(f"{name}: n={H.n} near={st['n_near']} far={st['n_far']} max_rank={st['max_rank']} "
          f"iters={s._last_solve_iters[0]} {dt:.1f} s", flush=True)


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(CASES)):
        record(name)
