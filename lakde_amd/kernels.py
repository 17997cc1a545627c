"""The reference's three native ops, with the reference's Python signatures,
running on the HIP library (lakde/kernels.py:93-120 is the interface mirrored).

    ind2ptr(ind, M, out=None)
    calc_W_inv(rowptr, col, value, Xn, Xm)
    chol_quad_form(W_triu, Xn, Xm)

They are also registered as ``torch.ops.lakde.{ind2ptr_cuda, W_inv_cuda,
chol_quad_form_cuda}`` with the reference schemas (csrc/cuda/convert_cuda.h:27,
csrc/cuda/W_inv_cuda.h:5-7, csrc/cuda/chol_quad_form.h:5-6), so code that calls
the ops directly keeps working.

Differences from the reference, all deliberate: CPU tensors raise (there is no
fallback path); unsupported dtypes raise instead of NaN-filling ``out``
(csrc/cuda/W_inv_cuda.cu:132-134, convert_cuda.cu:60-62); float16 inputs
are read into float32 arithmetic and the result is rounded once to float16
(the reference computes in half arithmetic); ``chol_quad_form`` reads only
the upper triangle of W_triu.
"""

import torch

from . import _lib
from ._lib import check, fp_code, idx_code, ptr, require_gpu, stream_of

__all__ = ["ind2ptr", "calc_W_inv", "chol_quad_form"]


def ind2ptr(ind, M, out=None):
    """CSR row pointer of sorted row indices (lakde/kernels.py:93-100)."""
    assert ind.is_contiguous()
    if out is None:
        out = ind.new_empty(M + 1)
    require_gpu(ind, out)
    if out.numel() != M + 1 or not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor of M + 1 elements")
    lib = _lib.load()
    check(lib.lakde_ind2ptr(ptr(ind), idx_code(ind.dtype), ind.numel(), int(M), ptr(out),
                            idx_code(out.dtype), stream_of(ind)), "ind2ptr")
    return out


def _calc_W_inv_into(rowptr, col, value, Xn, Xm, out):
    """Write the segmented outer-product sums into ``out`` (N, D, D) in place,
    with no other allocation (csrc/cuda/W_inv_cuda.h:5-7 semantics)."""
    assert rowptr.is_contiguous()
    assert col.is_contiguous()
    assert value.is_contiguous()
    assert Xn.is_contiguous()
    assert Xm.is_contiguous()
    require_gpu(rowptr, col, value, Xn, Xm, out)
    N, D = Xn.shape
    if rowptr.numel() != N + 1 or Xm.dim() != 2 or Xm.size(1) != D:
        raise ValueError("shape mismatch between rowptr, Xn and Xm")
    if not (value.dtype == Xn.dtype == Xm.dtype):
        raise TypeError("value, Xn and Xm must share one floating dtype")
    if tuple(out.shape) != (N, D, D) or not out.is_contiguous() or out.dtype != value.dtype:
        raise ValueError("out must be a contiguous (N, D, D) tensor of the value dtype")
    lib = _lib.load()
    check(lib.lakde_w_inv(ptr(rowptr), idx_code(rowptr.dtype), ptr(col), idx_code(col.dtype),
                          ptr(value), ptr(Xn), ptr(Xm), N, Xm.size(0), D, fp_code(value.dtype),
                          ptr(out), stream_of(value)), "calc_W_inv")
    return out


def calc_W_inv(rowptr, col, value, Xn, Xm):
    """out[n] = sum_e value[e] (Xm[col[e]] - Xn[n]) (...)^T over row n
    (lakde/kernels.py:102-113)."""
    N, D = Xn.shape
    return _calc_W_inv_into(rowptr, col, value, Xn, Xm, value.new_empty(N, D, D))


def _chol_quad_form_into(W_triu, Xn, Xm, out):
    require_gpu(W_triu, Xn, Xm, out)
    if W_triu.dim() != 3 or Xn.dim() != 2 or Xm.dim() != 2:
        raise ValueError("expected W_triu (N, D, D), Xn (N, D), Xm (M, D)")
    N, D = Xn.shape
    M = Xm.size(0)
    if Xm.size(1) != D or tuple(W_triu.shape) != (N, D, D) or tuple(out.shape) != (N, M):
        raise ValueError("shape mismatch in chol_quad_form")
    if not (W_triu.dtype == Xn.dtype == Xm.dtype == out.dtype):
        raise TypeError("all chol_quad_form tensors must share one floating dtype")
    lib = _lib.load()
    sW, sN, sM, sO = W_triu.stride(), Xn.stride(), Xm.stride(), out.stride()
    check(lib.lakde_chol_quad_form(ptr(W_triu), sW[0], sW[1], sW[2], ptr(Xn), sN[0], sN[1],
                                   ptr(Xm), sM[0], sM[1], ptr(out), sO[0], sO[1], N, M, D,
                                   fp_code(W_triu.dtype), stream_of(W_triu)), "chol_quad_form")
    return out


def chol_quad_form(W_triu, Xn, Xm):
    """Q[n, m] = ||W_triu[n]^T (Xm[m] - Xn[n])||^2 (lakde/kernels.py:115-120).
    Arbitrary strides (including stride-0 broadcasts) are honoured."""
    out = W_triu.new_empty(Xn.size(0), Xm.size(0))
    return _chol_quad_form_into(W_triu, Xn, Xm, out)


# ------------------------------------------------- torch.ops.lakde schemas
def _register_torch_ops():
    try:
        lib = torch.library.Library("lakde", "DEF")
    except RuntimeError:  # namespace already owned (e.g. imported twice)
        return None
    lib.define("ind2ptr_cuda(Tensor ind, int M, Tensor(a!) out) -> Tensor(a!)")
    lib.define("W_inv_cuda(Tensor rowptr, Tensor col, Tensor value, Tensor Xn, Tensor Xm, "
               "Tensor(a!) out) -> Tensor(a!)")
    lib.define("chol_quad_form_cuda(Tensor W_triu, Tensor Xn, Tensor Xm, Tensor(a!) out) "
               "-> Tensor(a!)")

    def _ind2ptr_impl(ind, M, out):
        return ind2ptr(ind, M, out=out)

    def _w_inv_impl(rowptr, col, value, Xn, Xm, out):
        return _calc_W_inv_into(rowptr, col, value, Xn, Xm, out)

    def _qf_impl(W_triu, Xn, Xm, out):
        return _chol_quad_form_into(W_triu, Xn, Xm, out)

    lib.impl("ind2ptr_cuda", _ind2ptr_impl, "CUDA")
    lib.impl("W_inv_cuda", _w_inv_impl, "CUDA")
    lib.impl("chol_quad_form_cuda", _qf_impl, "CUDA")
    return lib


_TORCH_LIB = _register_torch_ops()
