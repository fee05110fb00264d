"""mw_zgesv_noalloc (the factor-and-solve whose LU workspace is a static
device array) against mw_zgesv with a caller-allocated workspace: two calls
in a row, the second on a smaller system, then a size beyond the static
workspace refused with MW_ERANGE before anything is launched."""

import numpy as np
import pytest

from conftest import rel_err
from momwire_amd import _lib

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _system(n, p, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    M += 4 * np.sqrt(n) * np.eye(n) * (seed % 2)  # one dominant, one pivoting system
    B = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    return np.ascontiguousarray(np.hstack([M, B]))


def _solve(MB, n, p, noalloc):
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    dA = torch.from_numpy(MB).cuda()
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    if noalloc:
        _lib.check(lib.mw_zgesv_noalloc(n, p, dA.data_ptr(), n + p, ipiv.data_ptr(),
                                        info.data_ptr(), None, st), "zgesv_noalloc")
    else:
        work = torch.empty(int(lib.mw_getrf_workspace_bytes(n)), dtype=torch.uint8,
                           device="cuda")
        _lib.check(lib.mw_zgesv(n, p, dA.data_ptr(), n + p, ipiv.data_ptr(), work.data_ptr(),
                                info.data_ptr(), None, st), "zgesv")
    torch.cuda.synchronize()
    return dA.cpu().numpy(), ipiv.cpu().numpy(), int(info.cpu()[0])


def test_noalloc_gesv_matches_workspace_gesv():
    for n, p, seed in ((5000, 40, 1), (700, 9, 2)):
        _say("zgesv_noalloc vs zgesv", n, p)
        MB = _system(n, p, seed)
        A0, piv0, info0 = _solve(MB, n, p, noalloc=False)
        A1, piv1, info1 = _solve(MB, n, p, noalloc=True)
        assert info0 == 0 and info1 == 0
        # the same kernels on the same data (the tournament's merge order may
        # differ between runs, so bit equality is not required)
        assert rel_err(A1[:, n:], A0[:, n:]) < 1e-12
        X = A1[:, n:]
        ref = np.linalg.solve(MB[:, :n], MB[:, n:])
        assert rel_err(X, ref) < 1e-10


def test_noalloc_gesv_refuses_beyond_static_workspace():
    """n = 16385 is one past the static workspace's limit: refused on the
    host before any launch.  The buffers are sized for that n all the same
    (a zero matrix: a factorisation that did run would only report a zero
    pivot), so no outcome of the call can touch memory outside them."""
    _say("zgesv_noalloc size limit")
    lib = _lib.load()
    n, p = 16385, 1
    A = torch.zeros((n, n + p), dtype=torch.complex128, device="cuda")
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = lib.mw_zgesv_noalloc(n, p, A.data_ptr(), n + p, ipiv.data_ptr(), info.data_ptr(),
                              None, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == _lib.MW_ERANGE
    assert b"static scratch" in lib.mw_last_error()
    del A
