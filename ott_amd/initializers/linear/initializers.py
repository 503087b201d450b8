"""Sinkhorn initializers (initializers/linear/initializers.py:30-131 in OTT)."""

from __future__ import annotations

from typing import Any, Tuple

import torch

__all__ = ["SinkhornInitializer", "DefaultInitializer"]


class SinkhornInitializer:
  """Base class: ``init_fu`` / ``init_gv`` then -inf on zero weights (:70-99)."""

  def init_fu(self, ot_prob, lse_mode: bool, **kwargs: Any) -> torch.Tensor:
    raise NotImplementedError

  def init_gv(self, ot_prob, lse_mode: bool, **kwargs: Any) -> torch.Tensor:
    raise NotImplementedError

  def __call__(self, ot_prob, lse_mode: bool = True, **kwargs: Any
               ) -> Tuple[torch.Tensor, torch.Tensor]:
    fu = self.init_fu(ot_prob, lse_mode=lse_mode, **kwargs)
    gv = self.init_gv(ot_prob, lse_mode=lse_mode, **kwargs)
    mask_value = -float("inf") if lse_mode else 0.0
    fu = torch.where(ot_prob.a > 0.0, fu, torch.full_like(fu, mask_value))
    gv = torch.where(ot_prob.b > 0.0, gv, torch.full_like(gv, mask_value))
    return fu, gv


class DefaultInitializer(SinkhornInitializer):
  """Zero potentials (:111-131)."""

  def init_fu(self, ot_prob, lse_mode: bool, **kwargs: Any) -> torch.Tensor:
    a = ot_prob.a
    return torch.zeros_like(a) if lse_mode else torch.ones_like(a)

  def init_gv(self, ot_prob, lse_mode: bool, **kwargs: Any) -> torch.Tensor:
    b = ot_prob.b
    return torch.zeros_like(b) if lse_mode else torch.ones_like(b)
