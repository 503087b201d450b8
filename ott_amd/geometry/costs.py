"""Cost functions of point clouds.

Mirrors the translation-invariant costs of ``ott.geometry.costs``
(geometry/costs.py): ``c(x, y) = h(x - y)``.  The squared-Euclidean cost
(:547-573), the reference's default for PointCloud (geometry/pointcloud.py:68),
runs on the MFMA half-step through its cross term.  ``PNormP`` (:304-334),
``SqPNorm`` (:269-301), ``EuclideanP`` (:337-361) and ``Euclidean``
(:528-545) sum a per-feature term of ``z = x - y`` on the vector ALUs, online
(``ottx_problem.cost``); for p = 2 the first three are (half) the squared
Euclidean cost and take its MFMA path.  ``h`` evaluates a cost on torch
tensors, for checks and small host-side uses.
"""

from __future__ import annotations

from typing import Tuple

import torch

__all__ = ["CostFn", "TICost", "SqEuclidean", "PNormP", "SqPNorm", "EuclideanP", "Euclidean"]

# enum ottx_cost of include/ottx.h
COST_SQEUCLIDEAN = 0
COST_PNORMP = 1
COST_SQPNORM = 2
COST_EUCLIDEANP = 3
COST_EUCLIDEAN = 4


def _pnorm(z: torch.Tensor, p: float) -> torch.Tensor:
  return torch.sum(torch.abs(z) ** p, dim=-1) ** (1.0 / p)


class CostFn:
  """Base class of costs (geometry/costs.py:39-121)."""

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    raise NotImplementedError

  def kernel_spec(self) -> Tuple[int, float, float]:
    """(ottx cost, p, factor): the kernels evaluate ``factor * cost(kind, p)``;
    factor differs from 1 only where a cost is a multiple of the squared
    Euclidean one and runs as it."""
    raise NotImplementedError(f"{type(self).__name__} is not supported by the ott_amd kernels")


class TICost(CostFn):
  """Translation-invariant cost ``c(x, y) = h(x - y)`` (geometry/costs.py:125-160)."""

  def h(self, z: torch.Tensor) -> torch.Tensor:
    raise NotImplementedError

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    return self.h(x - y)


class SqEuclidean(TICost):
  """``c(x, y) = |x|^2 + |y|^2 - 2 <x, y>`` (geometry/costs.py:547-573)."""

  def norm(self, x: torch.Tensor) -> torch.Tensor:
    return torch.sum(x ** 2, dim=-1)

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    return self.norm(x) + self.norm(y) - 2.0 * torch.sum(x * y, dim=-1)

  def h(self, z: torch.Tensor) -> torch.Tensor:
    return torch.sum(z ** 2, dim=-1)

  def kernel_spec(self) -> Tuple[int, float, float]:
    return COST_SQEUCLIDEAN, 2.0, 1.0

  def __eq__(self, other) -> bool:
    return type(other) is SqEuclidean

  def __hash__(self) -> int:
    return hash("SqEuclidean")


class _PCost(TICost):

  def __init__(self, p: float):
    super().__init__()
    p = float(p)
    if not p >= 1.0:
      raise ValueError(f"p must be >= 1, got {p}")
    self.p = p
    self.q = 1.0 / (1.0 - 1.0 / p) if p > 1.0 else float("inf")

  def __eq__(self, other) -> bool:
    return type(other) is type(self) and other.p == self.p

  def __hash__(self) -> int:
    return hash((type(self).__name__, self.p))

  def __repr__(self) -> str:
    return f"{type(self).__name__}(p={self.p})"


class PNormP(_PCost):
  """``h(z) = ||z||_p^p / p`` (geometry/costs.py:304-334)."""

  def h(self, z: torch.Tensor) -> torch.Tensor:
    return torch.sum(torch.abs(z) ** self.p, dim=-1) / self.p

  def kernel_spec(self) -> Tuple[int, float, float]:
    if self.p == 2.0:
      return COST_SQEUCLIDEAN, 2.0, 0.5
    return COST_PNORMP, self.p, 1.0


class SqPNorm(_PCost):
  """``h(z) = ||z||_p^2 / 2`` (geometry/costs.py:269-301)."""

  def h(self, z: torch.Tensor) -> torch.Tensor:
    return 0.5 * _pnorm(z, self.p) ** 2

  def kernel_spec(self) -> Tuple[int, float, float]:
    if self.p == 2.0:
      return COST_SQEUCLIDEAN, 2.0, 0.5
    return COST_SQPNORM, self.p, 1.0


class EuclideanP(_PCost):
  """``h(z) = ||z||_2^p`` (geometry/costs.py:337-361)."""

  def h(self, z: torch.Tensor) -> torch.Tensor:
    return torch.sqrt(torch.sum(z ** 2, dim=-1)) ** self.p

  def kernel_spec(self) -> Tuple[int, float, float]:
    if self.p == 2.0:
      return COST_SQEUCLIDEAN, 2.0, 1.0
    return COST_EUCLIDEANP, self.p, 1.0


class Euclidean(CostFn):
  """``c(x, y) = ||x - y||_2`` (geometry/costs.py:528-545).  The reference
  does not cast it as a TICost (its h would not be strictly convex); it is
  still translation invariant, which is all the kernels use."""

  def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    return torch.sqrt(torch.sum((x - y) ** 2, dim=-1))

  def kernel_spec(self) -> Tuple[int, float, float]:
    return COST_EUCLIDEAN, 1.0, 1.0

  def __eq__(self, other) -> bool:
    return type(other) is Euclidean

  def __hash__(self) -> int:
    return hash("Euclidean")
