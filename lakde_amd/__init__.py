"""lakde_amd: lakde's hot path (Cholesky quadratic forms with a streaming
log-sum-exp, sparse covariance accumulation) as hand-written HIP kernels for
MI355X (gfx950), behind the reference's op signatures and model classes.

    from lakde_amd import LocalKNearestKDE, LocalHierarchicalKDE
"""

from .local_hierarchical_kde import LocalHierarchicalKDE
from .local_knn_kde import LocalKNearestKDE

__all__ = ["LocalKNearestKDE", "LocalHierarchicalKDE"]
