"""Selection of the learned warm start used by the benchmark.

Returns the initializer whose trained weights ship with the package for a
given feature dimension, or ``(None, "zeros")`` when none is present, in which
case the solver starts from zero potentials (OTT's DefaultInitializer).
"""

from __future__ import annotations

from typing import Callable, Optional, Tuple


def bench_initializer(d: int) -> Tuple[Optional[Callable], str]:
  return None, "zeros"
