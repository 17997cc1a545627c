"""CPU checks of the C-ABI library: it loads, exports every symbol declared in
include/lakde_amd.h, and the host-side argument validation and size queries
behave (no GPU work is launched here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lakde_amd.h")
LIB = os.path.join(ROOT, "lakde_amd", "liblakde_amd.so")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lakde_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("liblakde_amd.so not built (run python -m lakde_amd.build)")
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    names = _declared()
    for must in ("lakde_ind2ptr", "lakde_w_inv", "lakde_chol_quad_form", "lakde_estep",
                 "lakde_posterior", "lakde_rebalance", "lakde_pred_density"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_matches_header(lib):
    from lakde_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_size_queries(lib):
    lib.lakde_record_floats.restype = ctypes.c_int64
    lib.lakde_record_floats.argtypes = [ctypes.c_int64]
    for D in (1, 2, 6, 21, 43, 63):
        R = lib.lakde_record_floats(D)
        assert R % 4 == 0 and R >= D * (D + 1) // 2 + D + 4
    lib.lakde_estep_workspace.restype = ctypes.c_size_t
    lib.lakde_estep_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64]
    # one 64-bit ballot mask per (kernel, 64-point chunk) at least
    assert lib.lakde_estep_workspace(200_000, 512) >= 512 * 3125 * 8


def test_argument_validation_without_gpu(lib):
    """Invalid sizes and dtypes are rejected on the host before any launch."""
    from lakde_amd import _lib
    L = _lib.load()
    vp = ctypes.c_void_p
    rc = L.lakde_ind2ptr(vp(8), 0, 4, -1, vp(8), 0, vp(0))
    assert rc == 1 and b"negative" in L.lakde_last_error()
    rc = L.lakde_ind2ptr(vp(8), 7, 4, 3, vp(8), 0, vp(0))
    assert rc == 2
    rc = L.lakde_w_inv(vp(8), 0, vp(8), 0, vp(8), vp(8), vp(8), 3, 3, 600, 2, vp(8), vp(0))
    assert rc == 1
    rc = L.lakde_chol_quad_form(vp(8), 1, 1, 1, vp(8), 1, 1, vp(8), 1, 1, vp(8), 1, 1,
                                2, 2, 2, 9, vp(0))
    assert rc == 2
    args = _lib.EstepArgs()
    args.D = 64
    assert L.lakde_estep(ctypes.byref(args), vp(0)) == 1
    pargs = _lib.PosteriorArgs()
    pargs.D = 70
    assert L.lakde_posterior(ctypes.byref(pargs), vp(0)) == 1


def test_product_refuses_cpu_tensors():
    torch = pytest.importorskip("torch")
    from lakde_amd import kernels
    from lakde_amd._lib import LakdeError
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    with pytest.raises(LakdeError):
        kernels.ind2ptr(torch.zeros(3, dtype=torch.int32), 4)
    with pytest.raises(LakdeError):
        kernels.chol_quad_form(torch.eye(2)[None], torch.zeros(1, 2), torch.zeros(3, 2))


def test_knn_workspace_query_and_argument_checks(lib):
    """lakde_knn_graph: the workspace size query and the host-side checks of
    k (no launch happens for rejected arguments or empty inputs)."""
    lib.lakde_knn_graph_workspace.restype = ctypes.c_size_t
    lib.lakde_knn_graph_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64]
    # one split at N = 1M (>= 2048 blocks of 256 queries): N k 8-byte entries
    # plus N list sizes, each rounded to 256 bytes
    n, k = 1_000_000, 128
    assert lib.lakde_knn_graph_workspace(n, k) == (n * k * 8 + 255) // 256 * 256 + (n * 4 + 255) // 256 * 256
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.lakde_knn_graph.argtypes = [vp, i64, i64, i64, ctypes.c_int, vp, ctypes.c_int, vp, vp,
                                    ctypes.c_size_t, vp]
    lib.lakde_knn_graph.restype = ctypes.c_int
    dummy = ctypes.c_void_p(16)
    # k > N - 1 without self, D out of range, bad dtype: EINVAL before any launch
    assert lib.lakde_knn_graph(dummy, 10, 3, 10, 0, dummy, 0, None, dummy, 1 << 20, None) == 1
    assert lib.lakde_knn_graph(dummy, 10, 64, 3, 0, dummy, 0, None, dummy, 1 << 20, None) == 1
    assert lib.lakde_knn_graph(dummy, 10, 3, 3, 0, dummy, 7, None, dummy, 1 << 20, None) == 1
    # N = 0: nothing to do
    assert lib.lakde_knn_graph(dummy, 0, 3, 3, 0, dummy, 0, None, dummy, 0, None) == 0


def test_column_order_cpu():
    """utils.column_order (the pair-skipping E-step's column order): a
    permutation whose 64-column chunks are boxed exactly, the partial chunk
    last, and a k-d split on each node's widest coordinate (a chunk is much
    tighter than a random one)."""
    import numpy as np
    import torch
    from lakde_amd.utils import column_order
    rng = np.random.default_rng(4)
    X = torch.from_numpy((rng.standard_normal((5000, 4)) * [1, 3, 0.5, 2]).astype(np.float32))
    perm, box = column_order(X)
    p = perm.long()
    assert torch.equal(torch.sort(p).values, torch.arange(5000))
    nch = (5000 + 63) // 64
    assert box.shape == (nch, 2, 4)
    for c in range(nch):
        pts = X[p[c * 64:(c + 1) * 64]]
        assert torch.equal(box[c, 0], pts.amin(0)) and torch.equal(box[c, 1], pts.amax(0))
    widths = (box[:, 1] - box[:, 0]).log().sum(1).mean()
    rnd = X[torch.randperm(5000)[:4992]].view(78, 64, 4)
    rwidths = (rnd.amax(1) - rnd.amin(1)).log().sum(1).mean()
    assert widths < rwidths - 3.0
