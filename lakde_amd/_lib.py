"""ctypes binding of liblakde_amd.so (the C ABI declared in include/lakde_amd.h).

The library is loaded lazily on first use and the load fails loudly: there is
no CPU fallback anywhere in the product path.  Build it with
``python -m lakde_amd.build``.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LAKDE_AMD_LIB", os.path.join(_HERE, "liblakde_amd.so"))

LAKDE_I32, LAKDE_I64, LAKDE_F32, LAKDE_F64, LAKDE_F16 = 0, 1, 2, 3, 4
LAKDE_OK, LAKDE_EINVAL, LAKDE_ECAPACITY, LAKDE_ESTABILIZE = 0, 1, 4, 5
ESTEP_ACCUM, ESTEP_FIXED, ESTEP_UPDATE = 0, 1, 2
PRIOR_KNN, PRIOR_SHARED, PRIOR_KNN_CACHED = 0, 1, 2

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_int = ctypes.c_int
_f32 = ctypes.c_float
_sz = ctypes.c_size_t


class PosteriorArgs(ctypes.Structure):
    _fields_ = [
        ("nb", _i64), ("D", _i64), ("prior_mode", _int),
        ("prior", _vp), ("prior_jitter", _vp), ("S", _vp), ("nu", _vp), ("a", _vp), ("b", _vp),
        ("nu0", _f32), ("W_jitter", _vp), ("Xn", _vp), ("Linv", _vp), ("prior_tril", _vp),
        ("logdet_W", _vp), ("logdet_prior", _vp), ("trace", _vp), ("rec", _vp), ("status", _vp),
        ("b0", _f32), ("b_out", _vp),
    ]


class EstepArgs(ctypes.Structure):
    _fields_ = [
        ("N", _i64), ("D", _i64), ("nb", _i64), ("row0", _i64),
        ("X", _vp), ("rec", _vp), ("mode", _int), ("threshold", _f32),
        ("c_old", _vp), ("old_colsum", _vp), ("c", _vp),
        ("rowptr", _vp), ("out_row", _vp), ("out_col", _vp), ("out_val", _vp),
        ("cap", _i64), ("nnz_dev", _vp), ("workspace", _vp), ("workspace_bytes", _sz),
        ("lazy_G", _vp), ("lazy_unsafe", _vp), ("col_perm", _vp), ("chunk_box", _vp),
        ("lazy_factor", _vp),
    ]


# name -> (restype, argtypes)
class StepArgs(ctypes.Structure):
    _fields_ = [
        ("post", PosteriorArgs), ("estep", EstepArgs),
        ("knn_graph", _vp), ("knn_graph_dtype", _int), ("knn_k", _i64), ("nu0", _f32),
        ("block", _i64), ("num_blocks", _i64), ("cap", _i64),
        ("nnz_all", _vp), ("cols_all", _vp), ("vals_all", _vp),
        ("G", _vp), ("U", _vp), ("U_old", _vp), ("lazy_ws", _vp), ("rebal_ws", _vp),
        ("fold_first", _int), ("bf16_frag", _vp), ("host_flags", _vp), ("ev_main", _vp * 2),
        ("ev_flags", _vp), ("flag_words", _vp), ("abort_flag", _vp),
        ("screen_xb", _vp), ("screen_xnorm", _vp), ("screen_origin", _vp), ("fp32_mfma", _int),
        ("screen_split_ws", _vp), ("screen_split_bytes", _sz),
    ]


_SIGNATURES = {
    "lakde_version": (_int, []),
    "lakde_last_error": (ctypes.c_char_p, []),
    "lakde_ind2ptr": (_int, [_vp, _int, _i64, _i64, _vp, _int, _vp]),
    "lakde_w_inv": (_int, [_vp, _int, _vp, _int, _vp, _vp, _vp, _i64, _i64, _i64, _int, _vp, _vp]),
    "lakde_chol_quad_form": (_int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64,
                                    _vp, _i64, _i64, _i64, _i64, _i64, _int, _vp]),
    "lakde_record_floats": (_i64, [_i64]),
    "lakde_knn_prior_cov": (_int, [_vp, _i64, _i64, _i64, _i64, _vp, _int, _i64, _vp, _vp]),
    "lakde_posterior_contribs": (_int, [_vp, _int, _vp, _int, _vp, _vp, _vp, _vp, _i64, _i64,
                                        _i64, _f32, _vp, _vp, _vp]),
    "lakde_posterior": (_int, [ctypes.POINTER(PosteriorArgs), _vp]),
    "lakde_prior_cache_floats": (_i64, [_i64]),
    "lakde_knn_prior_cache": (_int, [_vp, _i64, _i64, _i64, _i64, _vp, _int, _i64, _vp, _vp, _vp,
                                     _vp, _vp, _vp]),
    "lakde_block_posterior": (_int, [ctypes.POINTER(PosteriorArgs), _vp, _i64, _i64, _vp, _vp, _vp,
                                     _vp, _vp, _int, _i64, _vp]),
    "lakde_colsum": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "lakde_estep_workspace": (_sz, [_i64, _i64]),
    "lakde_estep": (_int, [ctypes.POINTER(EstepArgs), _vp]),
    "lakde_estep_compact": (_int, [ctypes.POINTER(EstepArgs), _vp]),
    "lakde_estep_main": (_int, [ctypes.POINTER(EstepArgs), _vp]),
    "lakde_estep_finish": (_int, [ctypes.POINTER(EstepArgs), _vp]),
    "lakde_rebalance_workspace": (_sz, [_i64]),
    "lakde_rebalance": (_int, [_i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lakde_rebalance_lazy_begin": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lakde_rebalance_lazy_check": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "lakde_rebalance_lazy": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lakde_rebalance_fold": (_int, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lakde_flag_words_alloc": (_int, [ctypes.POINTER(_vp)]),
    "lakde_flag_words_free": (_int, [_vp]),
    "lakde_partial_step": (_int, [ctypes.POINTER(StepArgs), _vp]),
    "lakde_partial_step_resume": (_int, [ctypes.POINTER(StepArgs), _vp]),
    "lakde_partial_sweep": (_int, [ctypes.POINTER(StepArgs), _i64, _i64, ctypes.POINTER(_vp),
                                   ctypes.POINTER(_int), ctypes.POINTER(_int),
                                   ctypes.POINTER(_i64), _vp]),
    "lakde_pred_workspace": (_sz, [_i64, _i64]),
    "lakde_pred_density": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "lakde_pred_density_noscratch": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp]),
    "lakde_fill_f32": (_int, [_vp, _i64, _f32, _vp]),
    "lakde_knn_graph_workspace": (_sz, [_i64, _i64]),
    "lakde_knn_graph": (_int, [_vp, _i64, _i64, _i64, _int, _vp, _int, _vp, _vp, _sz, _vp]),
    "lakde_bf16_frag_bytes": (_i64, [_i64]),
    "lakde_pack_frag_bf16": (_int, [_vp, _i64, _i64, _vp, _vp]),
    "lakde_estep_main_bf16": (_int, [ctypes.POINTER(EstepArgs), _vp, _vp]),
    "lakde_m32_frag_bytes": (_i64, [_i64]),
    "lakde_pack_m32": (_int, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "lakde_estep_main_m32": (_int, [ctypes.POINTER(EstepArgs), _vp, _vp, _vp]),
    "lakde_pred_density_bf16": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "lakde_pred_screen_workspace": (_sz, [_i64, _i64]),
    "lakde_pred_density_screen": (_int, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _sz,
                                         _vp]),
    "lakde_screen_frag_bytes": (_i64, [_i64]),
    "lakde_screen_points_bytes": (_sz, [_i64, _i64]),
    "lakde_screen_points": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _vp]),
    "lakde_pack_screen": (_int, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "lakde_screen_knn_bound": (_int, [_vp, _vp, _i64, _i64, _vp, _int, _i64, _i64, _vp, _vp]),
    "lakde_estep_main_screen": (_int, [ctypes.POINTER(EstepArgs), _vp, _vp, _vp, _vp, _vp]),
    "lakde_screen_split_workspace": (_sz, [_i64, _i64]),
    "lakde_estep_main_screen_split": (_int, [ctypes.POINTER(EstepArgs), _vp, _vp, _vp, _vp, _sz,
                                             _vp]),
    "lakde_pack_records": (_int, [_vp, _vp, _vp, _i64, _i64, _vp, _vp]),
    "lakde_pred_backward_workspace": (_sz, [_i64, _i64, _i64]),
    "lakde_pred_backward": (_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _sz, _vp]),
}

EXPORTS = tuple(_SIGNATURES)

_lib = None


class LakdeError(RuntimeError):
    pass


def load(path=None):
    """Load (once) and return the ctypes library handle."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise LakdeError(
            f"liblakde_amd.so not found at {path}; build it with `python -m lakde_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().lakde_last_error().decode(errors="replace")
        raise LakdeError(f"{what} failed (status {rc}): {msg}")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def stream_of(t=None):
    """The current HIP stream of the tensor's device (torch's current stream)."""
    dev = t.device if t is not None else torch.device("cuda")
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def idx_code(dtype):
    if dtype == torch.int32:
        return LAKDE_I32
    if dtype == torch.int64:
        return LAKDE_I64
    raise TypeError(f"index tensors must be int32 or int64, got {dtype}")


def fp_code(dtype):
    if dtype == torch.float32:
        return LAKDE_F32
    if dtype == torch.float64:
        return LAKDE_F64
    if dtype == torch.float16:
        return LAKDE_F16
    raise TypeError(f"floating tensors must be float16, float32 or float64, got {dtype}")


def require_gpu(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise LakdeError("lakde_amd ops run on the GPU only; got a CPU tensor "
                             "(there is no CPU fallback)")
