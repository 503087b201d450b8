"""Epsilon scheduler (geometry/epsilon_scheduler.py:22-73 in OTT)."""

from __future__ import annotations

from typing import Optional

__all__ = ["Epsilon", "DEFAULT_EPSILON_SCALE"]

#: Scaling applied to the mean/std of the cost for the default epsilon.
DEFAULT_EPSILON_SCALE = 0.05


class Epsilon:
  """``eps_t = max(init * decay**t, 1) * target`` (epsilon_scheduler.py:47-60)."""

  def __init__(self, target: float, init: float = 1.0, decay: float = 1.0):
    assert decay <= 1.0, f"Decay must be <= 1, found {decay}."
    self.target = target
    self.init = init
    self.decay = decay

  def __call__(self, it: Optional[int]) -> float:
    if it is None:
      return self.target
    return max(self.init * (self.decay ** it), 1.0) * self.target

  @property
  def is_target_only(self) -> bool:
    """True when every iteration uses the target epsilon (``init <= 1``)."""
    return self.init <= 1.0

  def __repr__(self) -> str:
    return (f"{self.__class__.__name__}(target={float(self.target):.4f}, "
            f"init={self.init:.4f}, decay={self.decay:.4f})")
