"""lakde_pred_density_noscratch (the predictive density with its partial sums
in the library's static device scratch, no allocation) against
lakde_pred_density (caller workspace) through the C ABI: identical results
on two calls in a row, the second smaller, and a host-side rejection of a
problem whose partial sums exceed the scratch."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _records(rng, N, D):
    from lakde_amd.autograd import _pack_records
    A = rng.standard_normal((N, D, D)) * 0.3
    Winv = A @ np.swapaxes(A, 1, 2) + 0.2 * np.eye(D)
    Linv = np.tril(np.linalg.inv(np.linalg.cholesky(Winv)))
    Xc = rng.standard_normal((N, D))
    nu = D + 1 + rng.random(N) * 20
    f = lambda a: torch.tensor(a, dtype=torch.float32, device=DEV)  # noqa: E731
    return _pack_records(f(Linv), f(Xc), f(nu)), f(Xc)


def _with_workspace(lib, rec, Y, D):
    from lakde_amd._lib import check, ptr, stream_of
    N, M = rec.size(0), Y.size(0)
    nbytes = int(lib.lakde_pred_workspace(N, M))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=DEV)
    out = torch.empty(M, device=DEV)
    check(lib.lakde_pred_density(ptr(rec), N, ptr(Y), M, D, ptr(out), ptr(ws), nbytes,
                                 stream_of(Y)), "pred_density")
    return out


def _no_scratch(lib, rec, Y, D):
    from lakde_amd._lib import check, ptr, stream_of
    N, M = rec.size(0), Y.size(0)
    out = torch.full((M,), float("nan"), device=DEV)
    check(lib.lakde_pred_density_noscratch(ptr(rec), N, ptr(Y), M, D, ptr(out), stream_of(Y)),
          "pred_density_noscratch")
    return out


def test_pred_density_noscratch_matches_workspace_entry_point():
    from lakde_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(5)
    # two calls in a row, the second smaller (the scratch keeps the first
    # call's partial sums; nothing of them may leak into the second)
    for N, M, D in [(3000, 5000, 21), (700, 300, 21), (1500, 129, 6)]:
        print(f"noscratch: N={N} M={M} D={D}", flush=True)
        rec, _ = _records(rng, N, D)
        Y = torch.tensor(rng.standard_normal((M, D)), dtype=torch.float32, device=DEV)
        a = _no_scratch(lib, rec, Y, D)
        b = _with_workspace(lib, rec, Y, D)
        torch.cuda.synchronize()
        assert torch.isfinite(a).all()
        assert torch.equal(a, b), float((a - b).abs().max())


def test_pred_density_noscratch_rejects_oversized_problem():
    """Partial sums beyond the static scratch are refused on the host before
    any launch, with a message naming the sizes."""
    from lakde_amd import _lib
    lib = _lib.load()
    cap = 16 << 20
    N, M = 10_000, 3_000_000  # one split: 2 x 4 M bytes > 16 MiB
    assert lib.lakde_pred_workspace(N, M) > cap
    vp = ctypes.c_void_p(256)
    rc = lib.lakde_pred_density_noscratch(vp, N, vp, M, 8, vp, ctypes.c_void_p(0))
    assert rc == _lib.LAKDE_EINVAL
    assert b"static scratch" in lib.lakde_last_error()
