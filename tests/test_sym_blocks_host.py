"""CPU checks of the symmetric solve's block algebra (no GPU needed).

The HIP path (csrc/mw_linalg.hip: k_blkinv_init / k_blkinv_level / k_blkinv_vt
and sym_solve) forms, per 512 x 512 diagonal block of the factors, the
negated inverses N = -L^{-1} and V = -U^{-1} = N^T D^{-1} (U = D L^T) by
recursive doubling from 32 x 32 diagonal inverses, and solves the port
columns by one block product per 512 rows, skipping in the forward pass the
columns that have not started by a block's end.  These tests replay those
steps in numpy, in the kernels' order and signs, against numpy.linalg.inv and
numpy.linalg.solve.
"""

import numpy as np

NB, SYB = 32, 512


def _unit_lower(bs, rng, scale=0.15):
    L = np.tril(rng.standard_normal((bs, bs)) + 1j * rng.standard_normal((bs, bs)), -1) * scale
    return L + np.eye(bs)


def _doubling(L):
    """The kernels' steps on one block: N starts as the negated 32 x 32
    diagonal inverses; per level s, T = 0 - C N_P (kept apart, as in V's
    storage), then N's lower-left block = 0 - N_Q T."""
    bs = L.shape[0]
    N = np.zeros((SYB, SYB), dtype=complex)
    for p0 in range(0, bs, NB):
        p1 = min(p0 + NB, bs)
        N[p0:p1, p0:p1] = -np.linalg.inv(L[p0:p1, p0:p1])
    s = NB
    while s < SYB:
        for a0 in range(0, SYB, 2 * s):
            d0 = a0 + s
            if d0 >= bs:
                continue
            d1 = min(d0 + s, bs)
            T = 0 - L[d0:d1, a0:d0] @ N[a0:d0, a0:d0]
            N[d0:d1, a0:d0] = 0 - N[d0:d1, d0:d1] @ T
        s *= 2
    return N[:bs, :bs]


def test_block_inverse_by_doubling():
    rng = np.random.default_rng(5)
    for bs in (512, 404, 33):
        L = _unit_lower(bs, rng)
        N = _doubling(L)
        ref = np.linalg.inv(L)
        assert np.abs(N + ref).max() < 1e-12 * np.abs(ref).max()
        assert np.allclose(np.triu(N, 1), 0)
        # V = N^T D^{-1} is -(D L^T)^{-1}
        d = rng.standard_normal(bs) + 1j * rng.standard_normal(bs) + 3
        V = N.T / d[None, :]
        U = d[:, None] * L.T
        assert np.abs(V + np.linalg.inv(U)).max() < 1e-12 * np.abs(np.linalg.inv(U)).max()


def test_block_solve_with_first_rows():
    """sym_solve's forward and backward passes on a 3-block symmetric system
    (a partial last block), with a staircase of first rows: the forward pass
    works in block b on the columns up to the last one started by the block's
    end, and the answer equals the dense solve."""
    rng = np.random.default_rng(8)
    n, p = 2 * SYB + 300, 50
    L = _unit_lower(n, rng, scale=0.02)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n) + 4
    Z = (L * d[None, :]) @ L.T          # L D L^T, symmetric
    first = np.sort(rng.integers(0, n, p))
    first[:2] = 0
    B = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    B[np.arange(n)[:, None] < first[None, :]] = 0
    X = B.copy()
    nblk = (n + SYB - 1) // SYB
    Ns = [_doubling(L[b * SYB:min(n, (b + 1) * SYB), b * SYB:min(n, (b + 1) * SYB)])
          for b in range(nblk)]
    for b in range(nblk):
        K0, kend = b * SYB, min(n, (b + 1) * SYB)
        act = max([c + 1 for c in range(p) if first[c] < kend] + [0])
        assert np.all(X[K0:kend, act:] == 0)
        X[K0:kend, :act] = 0 - Ns[b] @ X[K0:kend, :act]
        X[kend:, :act] -= L[kend:, K0:kend] @ X[K0:kend, :act]
    U = d[:, None] * L.T
    for b in reversed(range(nblk)):
        K0, kend = b * SYB, min(n, (b + 1) * SYB)
        V = Ns[b].T / d[None, K0:kend]
        X[K0:kend] = 0 - V @ X[K0:kend]
        X[:K0] -= U[:K0, K0:kend] @ X[K0:kend]
    ref = np.linalg.solve(Z, B)
    assert np.abs(X - ref).max() < 1e-10 * np.abs(ref).max()
