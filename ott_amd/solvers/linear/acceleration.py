"""Momentum (solvers/linear/acceleration.py:127-172 in OTT).

Only the constant form (``start == 0``) runs on the GPU path; the adaptive
Lehmann momentum and Anderson acceleration raise ``NotImplementedError`` in
:class:`~ott_amd.solvers.linear.sinkhorn.Sinkhorn`.
"""

from __future__ import annotations

import dataclasses
import math

__all__ = ["Momentum", "AndersonAcceleration"]


@dataclasses.dataclass
class Momentum:
  start: int = 0
  error_threshold: float = math.inf
  value: float = 1.0
  inner_iterations: int = 1

  @property
  def is_constant(self) -> bool:
    return self.start == 0


@dataclasses.dataclass
class AndersonAcceleration:
  """Signature placeholder of acceleration.py:28-124 (not run on the GPU)."""
  memory: int = 2
  refresh_every: int = 1
  ridge_identity: float = 1e-2
