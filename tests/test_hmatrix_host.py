"""Host side of the block-low-rank operator: the block partition against the
partition the reference's HMatrixSolver built for the same structures
(tests/golden/hm_*.npz, hmatrix.py:939-980, _aca.py:61-134), and the
free-space-only refusal."""

import numpy as np
import pytest

from conftest import golden
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver
from momwire_amd._hpart import Partition

CASES = {
    "hm_array4x4": lambda: W.dipole_array(4, 4, 15, seed=2),
    "hm_chain120": lambda: W.wire_chain(120, 4, feed_every=10),
    "hm_array16x16": lambda: W.dipole_array(16, 16, 15, seed=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_partition_matches_reference(name):
    g = golden(name)
    mesh = BSplineSolver(**CASES[name]())._build()
    P = Partition(mesh, 1.0, 32)
    assert P.n == int(g["n"])
    assert np.array_equal(P.blocks, g["blocks"])
    assert np.array_equal(P.rows_pool, g["rows_pool"])
    assert np.array_equal(P.cols_pool, g["cols_pool"])
    assert P.n_far == g["aca_rank"].size


def test_partition_covers_z_once():
    mesh = BSplineSolver(**W.wire_chain(120, 4, feed_every=10))._build()
    P = Partition(mesh, 0.7, 16)
    cover = np.zeros((P.n, P.n), dtype=np.int32)
    for b in range(len(P.blocks)):
        cover[np.ix_(P.block_rows(b), P.block_cols(b))] += 1
    assert (cover == 1).all()
    leaves = np.concatenate(P.leaves)
    assert np.array_equal(np.sort(leaves), np.arange(P.n))
    assert max(len(l) for l in P.leaves) <= 16


def test_hmatrix_refuses_ground_and_loading():
    from momwire_amd.hmatrix import HMatrixSolver

    with pytest.raises(NotImplementedError, match="free space"):
        HMatrixSolver(**W.grounded_monopole("pec")).build_hmatrix()
    with pytest.raises(ValueError):
        HMatrixSolver(**W.straight_dipole(), aca_tol=0.0)
