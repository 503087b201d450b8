"""``linear.solve`` (solvers/linear/_solve.py:29-64 in OTT)."""

from __future__ import annotations

from typing import Any, Optional

from ott_amd.geometry.pointcloud import PointCloud
from ott_amd.problems.linear import linear_problem
from ott_amd.solvers.linear import sinkhorn

__all__ = ["solve"]


def solve(geom: PointCloud, a=None, b=None, tau_a: float = 1.0, tau_b: float = 1.0,
          rank: int = -1, **kwargs: Any) -> sinkhorn.SinkhornOutput:
  """Build ``LinearProblem(geom, a, b, tau_a, tau_b)`` and run ``Sinkhorn(**kwargs)``."""
  prob = linear_problem.LinearProblem(geom, a=a, b=b, tau_a=tau_a, tau_b=tau_b)
  if rank > 0:
    raise NotImplementedError("low-rank Sinkhorn (rank > 0) is outside ott_amd")
  solver = sinkhorn.Sinkhorn(**kwargs)
  return solver(prob)
