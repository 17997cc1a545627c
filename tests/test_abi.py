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
