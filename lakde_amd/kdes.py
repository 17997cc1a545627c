"""Fit driver shared by the models (lakde/kdes.py:8-136)."""

from abc import ABC, abstractmethod

import torch


class AbstractKDE(ABC):
    def __init__(self, block_size, verbose, logs):
        self.block_size = block_size
        self.verbose = verbose
        self.iter_steps = 0
        if logs:
            raise NotImplementedError("TensorBoard logging is not part of lakde_amd; pass logs=False")
        self.logger = None

    @abstractmethod
    def init_parameters_(self, X, sparsity_threshold=None):
        pass

    @abstractmethod
    def init_summaries_(self, X, sparsity_threshold=None):
        pass

    @abstractmethod
    def partial_step_(self, X, block_idx, sparsity_threshold=None):
        pass

    def fit(self, X, iterations=200, callbacks=None, sparsity_threshold=None, dataset_label=None):
        """One VI iteration = one sweep of partial steps over all kernel blocks
        (lakde/kdes.py:34-87).  ``callbacks=None`` runs no callback (the
        reference's default ELBO printer is observability only)."""
        N, D = X.shape
        assert X.is_contiguous(), "Training data must be contiguous, call .contiguous() before fitting"
        if sparsity_threshold is None:
            sparsity_threshold = 0
        num_blocks = (N - 1) // self.block_size + 1
        if self.iter_steps < 1:
            self.init_parameters_(X, sparsity_threshold)
            self.init_summaries_(X, sparsity_threshold)
            num_blocks = (N - 1) // self.block_size + 1
        try:
            for _ in range(iterations):
                for bi in range(num_blocks):
                    self.partial_step_(X, bi, sparsity_threshold=sparsity_threshold)
                    if callbacks is not None:
                        if isinstance(callbacks, (list, tuple)):
                            for cb in callbacks:
                                cb(X, self, self.iter_steps)
                        else:
                            callbacks(X, self, self.iter_steps)
                    self.iter_steps += 1
        except (KeyboardInterrupt, StopIteration):
            pass
        return self

    @abstractmethod
    def log_pred_density(self, X, Y):
        pass

    def pred_density(self, X, Y):
        return torch.exp_(self.log_pred_density(X, Y))

    def state_dict(self):
        return {"block_size": self.block_size, "verbose": self.verbose,
                "iter_steps": self.iter_steps}

    def load_state_dict(self, state_dict):
        self.block_size = state_dict["block_size"]
        self.verbose = state_dict["verbose"]
        self.iter_steps = state_dict["iter_steps"]

    def save(self, fname, state_dict=None):
        d = self.state_dict()
        if state_dict:
            d.update(state_dict)
        torch.save(d, fname)
