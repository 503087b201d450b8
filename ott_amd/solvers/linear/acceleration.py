"""Momentum (solvers/linear/acceleration.py:127-172 in OTT).

Constant momentum (``start == 0``) and the adaptive Lehmann momentum
(``start > 0``, acceleration.py:138-160) run on the GPU path, on balanced
and unbalanced problems and with any ``norm_error`` (the weight reads the
errors of the requested norm, as the reference records them); the adaptive
form needs ``start >= 2 inner_iterations`` and sequential updates
(:class:`~ott_amd.solvers.linear.sinkhorn.Sinkhorn` raises
``NotImplementedError`` otherwise).  Anderson acceleration is not run; its
class exists so that passing it raises the same error.
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
