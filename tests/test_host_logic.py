"""CPU tests of the host-side logic around the native calls: the SIGINT
deferral around model state updates (lakde/utils.py:240-255 semantics) and
the bookkeeping of the step driver's results.  The native library is
replaced by a stand-in, so no GPU work is launched."""

import ctypes
import os
import signal

import numpy as np
import pytest
import torch

from lakde_amd import _lib
from lakde_amd import local_adaptive_kde as lak
from lakde_amd.local_knn_kde import LocalKNearestKDE
from lakde_amd.utils import InterruptDelay


def test_interrupt_delay_defers_sigint():
    """A SIGINT inside the block does not interrupt it; KeyboardInterrupt is
    raised when the block exits, after the block's work is done."""
    done = []
    with pytest.raises(KeyboardInterrupt):
        with InterruptDelay():
            os.kill(os.getpid(), signal.SIGINT)
            done.append(1)  # still runs: the signal was only recorded
            done.append(2)
    assert done == [1, 2]
    assert signal.getsignal(signal.SIGINT) is signal.default_int_handler


def test_interrupt_delay_nested_and_quiet():
    """Nested blocks deliver one interrupt at the outermost exit; a block
    without a signal leaves the handler as it was."""
    with InterruptDelay():
        pass
    assert signal.getsignal(signal.SIGINT) is signal.default_int_handler
    order = []
    with pytest.raises(KeyboardInterrupt):
        with InterruptDelay():
            with InterruptDelay():
                os.kill(os.getpid(), signal.SIGINT)
                order.append("inner")
            order.append("outer")
    assert order == ["inner", "outer"]


class _FakeArena:
    def __init__(self, num_blocks):
        self.num_blocks = num_blocks
        self.nnz = [10] * num_blocks
        self.cap = 1 << 20


class _FakeLib:
    """lakde_partial_sweep stand-in: completes every requested step (entry
    count 100 + block), leaves the normalisers in the second buffer (cur =
    1), and delivers a SIGINT while 'inside' the native call."""

    def __init__(self, model):
        self.model = model
        self.calls = 0

    def lakde_partial_sweep(self, arr, first, count, cbuf, cur, dirty, done, stream):
        self.calls += 1
        flags = self.model._flags
        for b in range(first, first + count):
            flags[b] = 100 + b
        cur._obj.value = 1
        dirty._obj.value = 1
        done._obj.value = count
        os.kill(os.getpid(), signal.SIGINT)
        return _lib.LAKDE_OK


def _stub_model(monkeypatch, num_blocks=6):
    m = LocalKNearestKDE.__new__(LocalKNearestKDE)
    lak.LocallyAdaptiveKDE.__init__(m, None, 1e-6, 1e-6, 512, False, False)
    m.arena = _FakeArena(num_blocks)
    m._flags = np.zeros(num_blocks + 2, np.int32)
    m._active = 10.0 * num_blocks
    m._lazy = None
    m.rnm_log_consts = torch.zeros(8)
    m._c_new = torch.ones(8)
    m._step_cache = {"arr": (_lib.StepArgs * num_blocks)()}
    monkeypatch.setattr(m, "_step_args", lambda X, b: m._step_cache["arr"][b])
    fake = _FakeLib(m)
    monkeypatch.setattr(lak._lib, "load", lambda: fake)
    monkeypatch.setattr(lak, "stream_of", lambda X: ctypes.c_void_p(0))
    return m, fake


def test_sigint_during_native_sweep_keeps_state_consistent(monkeypatch):
    """A SIGINT that arrives while the native sweep runs is delivered only
    after the host bookkeeping of the completed steps: entry counts updated,
    the normaliser buffers swapped (cur = 1), the steps counted."""
    m, fake = _stub_model(monkeypatch)
    c0, c1 = m.rnm_log_consts, m._c_new
    with pytest.raises(KeyboardInterrupt):
        m._run_steps(torch.zeros(8, 2), 0, 6, 0.0, count_steps=True)
    assert fake.calls == 1
    assert m.arena.nnz == [100 + b for b in range(6)]
    assert m._active == float(sum(100 + b for b in range(6)))
    assert m.rnm_log_consts is c1 and m._c_new is c0
    assert m.iter_steps == 6


def test_fit_returns_consistent_model_after_sigint(monkeypatch):
    """fit() in chunks of sweep_chunk partial steps: the interrupt stops the
    fit after the native call in progress; fit swallows it (as the
    reference's fit does) and returns a model whose step count and entry
    counts agree with the steps the native side completed."""
    m, fake = _stub_model(monkeypatch, num_blocks=6)
    m.sweep_chunk = 4
    m.iter_steps = 1  # initialised: fit() continues without init
    monkeypatch.setattr(m, "materialize", lambda: None)
    X = torch.zeros(3000, 2)
    m.fit(X, iterations=3)
    assert fake.calls == 1
    assert m.iter_steps == 1 + 4
    assert m.arena.nnz[:4] == [100, 101, 102, 103] and m.arena.nnz[4:] == [10, 10]
    assert signal.getsignal(signal.SIGINT) is signal.default_int_handler
