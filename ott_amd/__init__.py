"""ott_amd: MI355X-native log-domain Sinkhorn on point clouds.

A drop-in for ``ott.solvers.linear.solve(PointCloud(x, y), a, b)`` backed by
hand-written HIP kernels for gfx950 (``ott_amd/csrc``, C ABI in
``include/ottx.h``).  Package layout mirrors ``ott``: ``geometry``,
``problems``, ``solvers``, ``initializers``, ``tools``.
"""

__version__ = "0.1.0"

from ott_amd import geometry, initializers, problems, solvers  # noqa: E402

__all__ = ["geometry", "initializers", "problems", "solvers", "__version__"]
