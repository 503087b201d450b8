"""Linear OT problem (problems/linear/linear_problem.py:30-205 in OTT)."""

from __future__ import annotations

from typing import Optional

import torch

from ott_amd.geometry.pointcloud import PointCloud, as_device_f32

__all__ = ["LinearProblem"]


class LinearProblem:
  """Geometry plus marginals and unbalancedness (linear_problem.py:51-63)."""

  def __init__(self, geom: PointCloud, a=None, b=None, tau_a: float = 1.0,
               tau_b: float = 1.0):
    self.geom = geom
    self._a = None if a is None else as_device_f32(a, geom.device)
    self._b = None if b is None else as_device_f32(b, geom.device)
    n, m = geom.shape
    if self._a is not None and self._a.shape != (n,):
      raise ValueError(f"a must have shape ({n},), got {tuple(self._a.shape)}")
    if self._b is not None and self._b.shape != (m,):
      raise ValueError(f"b must have shape ({m},), got {tuple(self._b.shape)}")
    self.tau_a = float(tau_a)
    self.tau_b = float(tau_b)

  @property
  def a(self) -> torch.Tensor:
    """linear_problem.py:65-71: uniform when not given."""
    if self._a is not None:
      return self._a
    n, _ = self.geom.shape
    return torch.full((n,), 1.0 / n, dtype=torch.float32, device=self.geom.device)

  @property
  def b(self) -> torch.Tensor:
    """linear_problem.py:73-79."""
    if self._b is not None:
      return self._b
    _, m = self.geom.shape
    return torch.full((m,), 1.0 / m, dtype=torch.float32, device=self.geom.device)

  @property
  def is_balanced(self) -> bool:
    return self.tau_a == 1.0 and self.tau_b == 1.0

  @property
  def is_uniform(self) -> bool:
    return self._a is None and self._b is None

  @property
  def is_equal_size(self) -> bool:
    return self.geom.shape[0] == self.geom.shape[1]

  @property
  def is_assignment(self) -> bool:
    return self.is_equal_size and self.is_uniform and self.is_balanced

  @property
  def epsilon(self) -> float:
    return self.geom.epsilon

  @property
  def dtype(self) -> torch.dtype:
    return self.geom.dtype
