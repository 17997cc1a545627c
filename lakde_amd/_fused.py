"""Thin wrappers of the fused entry points and the block arena that holds the
responsibilities on the device."""

import ctypes
import math

import torch

from . import _lib
from ._lib import ESTEP_ACCUM, EstepArgs, PosteriorArgs, check, idx_code, ptr, stream_of


def record_floats(D):
    return int(_lib.load().lakde_record_floats(D))


class Arena:
    """Responsibilities r_nm of all kernel blocks, in one device allocation.

    Block b keeps its entries at [b*cap, b*cap + nnz[b]) in canonical CSR
    order: rows local to the block (ascending), global columns ascending
    within a row.  This is the data of the reference's per-block
    ``rnm_blocks[b] = (indices (2, nnz) int32, values)``
    (lakde/local_adaptive_kde.py:145-153) with the row pointer kept alongside.
    ``nnz_dev`` carries three extra slots: the posterior status flag and the
    lazy rebalance's unsafe-step flag, read with the entry counts in one
    copy, and the step driver's abort word (set when a partial step needs
    the host, so a posterior of the next step enqueued ahead does nothing).
    """

    def __init__(self, num_blocks, block_size, cap, device):
        self.num_blocks = num_blocks
        self.block_size = block_size
        self.cap = int(cap)
        self.device = device
        i32 = torch.int32
        self.row = torch.zeros(num_blocks, self.cap, dtype=i32, device=device)
        self.col = torch.zeros(num_blocks, self.cap, dtype=i32, device=device)
        self.val = torch.zeros(num_blocks, self.cap, dtype=torch.float32, device=device)
        self.rowptr = torch.zeros(num_blocks, block_size + 1, dtype=i32, device=device)
        self.nnz_dev = torch.zeros(num_blocks + 3, dtype=i32, device=device)
        self.nnz = [0] * num_blocks

    @property
    def status_ptr(self):
        return self.nnz_dev[self.num_blocks:self.num_blocks + 1]

    @property
    def unsafe_ptr(self):
        return self.nnz_dev[self.num_blocks + 1:self.num_blocks + 2]

    @property
    def abort_ptr(self):
        return self.nnz_dev[self.num_blocks + 2:]

    def grow(self, need):
        new_cap = max(int(need * 1.25), int(self.cap * 1.5)) + 255
        new_cap = new_cap // 256 * 256
        for name in ("row", "col", "val"):
            old = getattr(self, name)
            new = torch.zeros(self.num_blocks, new_cap, dtype=old.dtype, device=old.device)
            new[:, :self.cap] = old
            setattr(self, name, new)
        self.cap = new_cap

    def block(self, b):
        """(indices (2, nnz) int32, values) views of block b."""
        n = self.nnz[b]
        return torch.stack([self.row[b, :n], self.col[b, :n]]), self.val[b, :n]

    def total_nnz(self):
        return sum(self.nnz)


class FlagWords:
    """The step driver's host-coherent flag words (lakde_flag_words_alloc):
    the row-pointer kernel writes each partial step's entry count and flags
    there, so the host waits for them without a copy or an event."""

    def __init__(self):
        self._lib = _lib.load()
        self._vp = ctypes.c_void_p  # kept: module globals may be gone at interpreter exit
        p = ctypes.c_void_p()
        check(self._lib.lakde_flag_words_alloc(ctypes.byref(p)), "flag_words_alloc")
        self.ptr = p.value

    def __del__(self):
        ptr_ = getattr(self, "ptr", None)
        if ptr_:
            self._lib.lakde_flag_words_free(self._vp(ptr_))
            self.ptr = None


def knn_prior_cov(X, row0, graph_blk, out):
    N, D = X.shape
    nb, K = graph_blk.shape
    check(_lib.load().lakde_knn_prior_cov(ptr(X), N, D, row0, nb, ptr(graph_blk),
                                          idx_code(graph_blk.dtype), K, ptr(out), stream_of(X)),
          "knn_prior_cov")
    return out


def prior_cache_floats(D):
    return int(_lib.load().lakde_prior_cache_floats(D))


def knn_prior_cache(X, row0, graph_blk, scratch, prior_jitter, logdet_prior, cache, status):
    """The kNN prior of rows row0.. factored once into cache (one entry per
    row: prior + j0 I and its factor, packed)."""
    N, D = X.shape
    nb, K = graph_blk.shape
    check(_lib.load().lakde_knn_prior_cache(ptr(X), N, D, row0, nb, ptr(graph_blk),
                                            idx_code(graph_blk.dtype), K, ptr(scratch),
                                            ptr(prior_jitter), ptr(logdet_prior), ptr(cache),
                                            ptr(status), stream_of(X)),
          "knn_prior_cache")


def posterior_contribs(rowptr, col, val, Xn, X, nu0, S, nu, colscale=None):
    nb, D = Xn.shape
    check(_lib.load().lakde_posterior_contribs(ptr(rowptr), idx_code(rowptr.dtype), ptr(col),
                                               idx_code(col.dtype), ptr(val), ptr(colscale),
                                               ptr(Xn), ptr(X),
                                               nb, X.size(0), D, float(nu0), ptr(S), ptr(nu),
                                               stream_of(X)),
          "posterior_contribs")


def posterior_args(prior_mode, prior, prior_jitter, S, nu, a, b, nu0, W_jitter, Xn, Linv,
                   prior_tril, logdet_W, logdet_prior, trace, rec, status, b0=0.0, b_out=None):
    nb, D = Xn.shape
    return PosteriorArgs(nb, D, prior_mode, ptr(prior), ptr(prior_jitter), ptr(S), ptr(nu),
                         ptr(a), ptr(b), float(nu0), ptr(W_jitter), ptr(Xn), ptr(Linv),
                         ptr(prior_tril), ptr(logdet_W), ptr(logdet_prior), ptr(trace), ptr(rec),
                         ptr(status), float(b0), ptr(b_out))


def posterior(*args, **kw):
    a = posterior_args(*args, **kw)
    check(_lib.load().lakde_posterior(a, stream_of(args[9])), "posterior")


def block_posterior(X, row0, rowptr, col, val, colscale, graph_blk, post_args):
    """Prior (kNN graph rows, or None for a shared prior), contributions and
    factorisation of one block from one call (lakde_block_posterior)."""
    N = X.size(0)
    gk = graph_blk.size(1) if graph_blk is not None else 0
    gdt = idx_code(graph_blk.dtype) if graph_blk is not None else 0
    check(_lib.load().lakde_block_posterior(post_args, ptr(X), N, row0, ptr(rowptr), ptr(col),
                                            ptr(val), ptr(colscale), ptr(graph_blk), gdt, gk,
                                            stream_of(X)),
          "block_posterior")


def colsum(col, val, nnz_dev, N, out, ws=None):
    """out[m] = sum of the block's values in column m (an order-independent
    sum; ws: N zeroed 16-byte accumulators, left zero)."""
    if ws is None:
        ws = fx_zeros(N, out.device)
    check(_lib.load().lakde_colsum(ptr(col), ptr(val), ptr(nnz_dev), N, ptr(out), ptr(ws),
                                   stream_of(val)),
          "colsum")
    return out


def fx_zeros(n, device):
    """n zeroed lakde_fx128 accumulators (order-independent column sums)."""
    return torch.zeros(n, 2, dtype=torch.int64, device=device)


def rebalance_workspace(N, device):
    """The exact rebalance's workspace: zeroed (its column sums must read
    zero on first use) and left so."""
    nbytes = int(_lib.load().lakde_rebalance_workspace(N))
    return torch.zeros((nbytes + 15) // 16, 2, dtype=torch.int64, device=device)


class EstepWorkspace:
    def __init__(self, N, nb, device):
        nbytes = int(_lib.load().lakde_estep_workspace(N, nb))
        # zero-filled: the E-step's per-segment counters start (and end) at zero
        self.buf = torch.zeros(nbytes, dtype=torch.uint8, device=device)


def estep_args(X, rec_blk, row0, mode, threshold, c, ws, c_old=None, old_colsum=None,
               arena=None, b=None, order=None):
    """``order``: None, or (col_perm, chunk_box) of utils.column_order — the
    column order of the pair-skipping E-step."""
    N, D = X.shape
    nb = rec_blk.size(0)
    if arena is not None:
        out = (ptr(arena.rowptr[b]), ptr(arena.row[b]), ptr(arena.col[b]), ptr(arena.val[b]),
               arena.cap, ptr(arena.nnz_dev[b:b + 1]))
    else:
        out = (ptr(None), ptr(None), ptr(None), ptr(None), 0, ptr(None))
    perm, box = order if order is not None else (None, None)
    return EstepArgs(N, D, nb, row0, ptr(X), ptr(rec_blk), mode, float(threshold or 0.0),
                     ptr(c_old), ptr(old_colsum), ptr(c), out[0], out[1], out[2], out[3], out[4],
                     out[5], ptr(ws.buf), ws.buf.numel(), None, None, ptr(perm), ptr(box))


def estep(X, rec_blk, row0, mode, threshold, c, ws, c_old=None, old_colsum=None,
          arena=None, b=None, compact_only=False, bf16=None, order=None, screen=None, m32=None,
          split=None):
    """One E-step launch sequence.  ``bf16=frag`` (the block's
    lakde_pack_frag_bf16 blocks) runs the main (quadratic form / LSE /
    ballot) kernel on the matrix cores in bf16x3; ``screen=(frag, points)``
    (the block's lakde_pack_screen blocks and the fit's ScreenPoints) runs
    the screening E-step, in its split form with ``split`` (a uint8 tensor
    of lakde_screen_split_workspace bytes); ``m32=(frag, points)``
    (lakde_pack_m32 blocks) the plain FP32 E-step on the matrix cores."""
    args = estep_args(X, rec_blk, row0, mode, threshold, c, ws, c_old, old_colsum, arena, b,
                      order)
    lib = _lib.load()
    st = stream_of(X)
    if compact_only:
        check(lib.lakde_estep_compact(args, st), "estep_compact")
        return
    if m32 is not None:
        frag, pts = m32
        check(lib.lakde_estep_main_m32(args, ptr(frag), ptr(pts.xnorm), st), "estep_main_m32")
    elif screen is not None and split is not None:
        frag, pts = screen
        check(lib.lakde_estep_main_screen_split(args, ptr(frag), ptr(pts.xb), ptr(pts.xnorm),
                                                ptr(split), split.numel(), st),
              "estep_main_screen_split")
    elif screen is not None:
        frag, pts = screen
        lb = pts.lbound if mode == ESTEP_ACCUM else None
        check(lib.lakde_estep_main_screen(args, ptr(frag), ptr(pts.xb), ptr(pts.xnorm), ptr(lb),
                                          st), "estep_main_screen")
    elif bf16 is not None:
        check(lib.lakde_estep_main_bf16(args, ptr(bf16), st), "estep_main_bf16")
    else:
        check(lib.lakde_estep_main(args, st), "estep_main")
    check(lib.lakde_estep_finish(args, st), "estep_finish")


def screen_frag_bytes(D):
    return int(_lib.load().lakde_screen_frag_bytes(D))


def screen_split_workspace(N, nb, device):
    n = int(_lib.load().lakde_screen_split_workspace(N, nb))
    return torch.empty(max(n, 1), dtype=torch.uint8, device=device)


class ScreenPoints:
    """The screening E-step's per-fit operands: the origin o (the mean of
    X, or the given one: test points are prepared with the origin of the
    training points the fragments were packed with), bf16(x_m - o) zero
    padded per point, and per point |x_m - o| then the norm of its bf16
    rounding residual (2N floats, rounded up)."""

    def __init__(self, X, origin=None):
        N, D = X.shape
        lib = _lib.load()
        self.origin = X.mean(0).contiguous() if origin is None else origin.contiguous()
        self.xb = torch.empty(max(int(lib.lakde_screen_points_bytes(N, D)), 1), dtype=torch.uint8,
                              device=X.device)
        self.xnorm = torch.empty(2 * N, device=X.device)  # |x - o|, then the bf16 residual norms
        check(lib.lakde_screen_points(ptr(X), N, D, ptr(self.origin), ptr(self.xb),
                                      ptr(self.xnorm), stream_of(X)), "screen_points")
        self.lbound = None  # lower bounds of the normalisers for ACCUM passes

    def knn_bound(self, rec, X, graph, J=16):
        """lbound[m] = the largest exact log rho of x_m under the kernels of
        its first J graph neighbours (a lower bound of its normaliser)."""
        N, D = X.shape
        self.lbound = torch.empty(N, device=X.device)
        check(_lib.load().lakde_screen_knn_bound(ptr(rec), ptr(X), N, D, ptr(graph),
                                                 idx_code(graph.dtype), graph.size(1), J,
                                                 ptr(self.lbound), stream_of(X)),
              "screen_knn_bound")
        return self.lbound


def m32_frag_bytes(D):
    return int(_lib.load().lakde_m32_frag_bytes(D))


def pack_m32(rec, D, origin, out):
    nb = rec.size(0)
    check(_lib.load().lakde_pack_m32(ptr(rec), nb, D, ptr(origin), ptr(out), stream_of(rec)),
          "pack_m32")
    return out


def pack_screen(rec, D, origin, out):
    nb = rec.size(0)
    check(_lib.load().lakde_pack_screen(ptr(rec), nb, D, ptr(origin), ptr(out), stream_of(rec)),
          "pack_screen")
    return out


def rebalance(arena, exclude_block, c_prev, c, ws):
    N = c.numel()
    check(_lib.load().lakde_rebalance(N, arena.num_blocks, arena.cap, exclude_block,
                                      ptr(arena.nnz_dev), ptr(arena.col), ptr(arena.val),
                                      ptr(c_prev), ptr(c), ptr(ws), stream_of(c)),
          "rebalance")


class LazyRebalance:
    """Column state of the lazy rebalance: the arena holds stored values u
    whose true values are u * G[col]; U are the stored column totals
    (order-independent fixed-point sums, N x 2 int64 = lakde_fx128).
    ``old_colsum`` holds the true old column sums of the updated block (read
    by the E-step); ``U_old`` and the workspace (3 N accumulators) stay zero
    between steps; ``exact_ws`` is the exact rebalance's workspace for the
    steps the stored form cannot represent."""

    def __init__(self, N, device):
        self.N = N
        self.G = torch.ones(N, device=device)
        self.U = fx_zeros(N, device)
        self.U_old = fx_zeros(N, device)
        self.old_colsum = torch.zeros(N, device=device)
        self.ws = fx_zeros(3 * N, device)
        self.exact_ws = rebalance_workspace(N, device)
        self.dirty = False  # G != 1 somewhere

    def end(self, arena, b, c_prev, c):
        if b is None:
            blk = (ptr(None), ptr(None), ptr(None))
        else:
            blk = (ptr(arena.col[b]), ptr(arena.val[b]), ptr(arena.nnz_dev[b:b + 1]))
        check(_lib.load().lakde_rebalance_lazy(
            self.N, blk[0], blk[1], blk[2], ptr(c_prev), ptr(c), ptr(self.G), ptr(self.U),
            ptr(self.U_old), ptr(self.ws), stream_of(c)),
            "rebalance_lazy")
        self.dirty = True

    def fold(self, arena):
        check(_lib.load().lakde_rebalance_fold(
            self.N, arena.num_blocks, arena.cap, ptr(arena.nnz_dev), ptr(arena.col),
            ptr(arena.val), ptr(self.G), ptr(self.U), stream_of(self.G)),
            "rebalance_fold")
        self.dirty = False


def pred_density(rec, Y, out):
    N = rec.size(0)
    M, D = Y.shape
    lib = _lib.load()
    nbytes = int(lib.lakde_pred_workspace(N, M))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Y.device)
    check(lib.lakde_pred_density(ptr(rec), N, ptr(Y), M, D, ptr(out), ptr(ws), nbytes,
                                 stream_of(Y)),
          "pred_density")
    return out


def knn_graph(X, k, include_self=False, dtype=torch.int32, return_dists=False):
    """k nearest neighbours of every row of X (N x D float32 on the device),
    rows sorted by distance with equal distances by index
    (lakde/utils.py:77-116)."""
    X = X.contiguous()
    if X.dtype != torch.float32:
        raise TypeError("knn_graph: X must be float32")
    N, D = X.shape
    out = torch.empty(N, k, dtype=dtype, device=X.device)
    dist = torch.empty(N, k, dtype=torch.float32, device=X.device) if return_dists else None
    lib = _lib.load()
    nbytes = int(lib.lakde_knn_graph_workspace(N, k))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=X.device)
    check(lib.lakde_knn_graph(ptr(X), N, D, k, int(bool(include_self)), ptr(out), idx_code(dtype),
                              ptr(dist) if dist is not None else None, ptr(ws), nbytes,
                              stream_of(X)),
          "knn_graph")
    return (dist, out) if return_dists else out


def bf16_frag_bytes(D):
    return int(_lib.load().lakde_bf16_frag_bytes(D))


def bf16_kp(D):
    return (D + 15) // 16 * 16


def pack_frag_bf16(rec, D, out):
    nb = rec.size(0)
    check(_lib.load().lakde_pack_frag_bf16(ptr(rec), nb, D, ptr(out), stream_of(rec)),
          "pack_frag_bf16")
    return out


def pred_density_screen(rec, frag, Y, pts, out):
    """The precision form's predictive density with the BF16 screen (frag:
    pack_screen of all N kernels; pts: ScreenPoints of Y with the fit's
    origin)."""
    N = rec.size(0)
    M, D = Y.shape
    lib = _lib.load()
    nbytes = int(lib.lakde_pred_screen_workspace(N, M))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Y.device)
    check(lib.lakde_pred_density_screen(ptr(rec), ptr(frag), N, ptr(Y), M, D, ptr(pts.xb),
                                        ptr(pts.xnorm), ptr(out), ptr(ws), nbytes, stream_of(Y)),
          "pred_density_screen")
    return out


def pred_density_bf16(rec, frag, Y, out):
    N = rec.size(0)
    M, D = Y.shape
    lib = _lib.load()
    nbytes = int(lib.lakde_pred_workspace(N, M))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=Y.device)
    check(lib.lakde_pred_density_bf16(ptr(rec), ptr(frag), N, ptr(Y), M, D, ptr(out), ptr(ws),
                                      nbytes, stream_of(Y)),
          "pred_density_bf16")
    return out


def fill_(t, v):
    check(_lib.load().lakde_fill_f32(ptr(t), t.numel(), float(v), stream_of(t)), "fill")
    return t


NEG_INF = -math.inf
