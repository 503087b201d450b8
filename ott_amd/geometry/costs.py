"""Cost functions on the accelerated path.

Only the squared-Euclidean cost (geometry/costs.py:547-572 in OTT) is
accelerated; it is also the reference's default cost of PointCloud
(geometry/pointcloud.py:68).
"""

from __future__ import annotations

import torch

__all__ = ["CostFn", "SqEuclidean"]


class CostFn:
  """Base class of costs (geometry/costs.py:39-121)."""

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    raise NotImplementedError


class SqEuclidean(CostFn):
  """``c(x, y) = |x|^2 + |y|^2 - 2 <x, y>`` (geometry/costs.py:547-561)."""

  def norm(self, x: torch.Tensor) -> torch.Tensor:
    return torch.sum(x ** 2, dim=-1)

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    return self.norm(x) + self.norm(y) - 2.0 * torch.sum(x * y, dim=-1)

  def __eq__(self, other) -> bool:
    return type(other) is SqEuclidean

  def __hash__(self) -> int:
    return hash("SqEuclidean")
