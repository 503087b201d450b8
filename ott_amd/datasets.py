"""Seeded synthetic point-cloud problems of the benchmark configurations.

Plays the role of ``ott.datasets`` for this package: every generator uses
``numpy.random.default_rng(seed)`` in float64 and returns float32 arrays, so
the same inputs can be handed to the GPU solver and to a CPU oracle.
"""

from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

__all__ = ["Problem", "uniform_shifted", "gaussian_shifted", "gaussian_mixture",
           "mixture_pair", "batched_uneven"]


@dataclasses.dataclass
class Problem:
  x: np.ndarray
  y: np.ndarray
  a: Optional[np.ndarray]
  b: Optional[np.ndarray]
  eps_mult: float              # epsilon = eps_mult * mean cost
  tau_a: float = 1.0
  tau_b: float = 1.0
  name: str = ""


def _normalise(w: np.ndarray) -> np.ndarray:
  return (w / w.sum()).astype(np.float32)


def uniform_shifted(n: int = 1024, m: int = 1024, d: int = 2, seed: int = 0) -> Problem:
  """x ~ U[0,1]^d, y ~ U[0,1]^d + 0.25, a, b = U(0,1) + 0.2 normalised; eps = 0.05 mean C."""
  rng = np.random.default_rng(seed)
  x = rng.uniform(size=(n, d))
  y = rng.uniform(size=(m, d)) + 0.25
  a = rng.uniform(size=n) + 0.2
  b = rng.uniform(size=m) + 0.2
  return Problem(x.astype(np.float32), y.astype(np.float32), _normalise(a), _normalise(b),
                 0.05, name=f"uniform_shifted_n{n}_m{m}_d{d}")


def gaussian_shifted(n: int = 16384, m: int = 16384, d: int = 64, seed: int = 1,
                     shift: float = 0.3) -> Problem:
  """x ~ N(0, I/d), y ~ N(shift / sqrt(d) * 1, I/d), uniform weights; eps = 0.05 mean C."""
  rng = np.random.default_rng(seed)
  x = rng.normal(size=(n, d)) / np.sqrt(d)
  y = rng.normal(size=(m, d)) / np.sqrt(d) + shift / np.sqrt(d)
  return Problem(x.astype(np.float32), y.astype(np.float32), None, None, 0.05,
                 name=f"gaussian_shifted_n{n}_m{m}_d{d}")


def gaussian_mixture(rng: np.random.Generator, n: int, d: int, k: int) -> np.ndarray:
  """k components, means ~ N(0, 4I/d), covariance I/d, weights ~ Dirichlet(1_k)."""
  means = rng.normal(size=(k, d)) * (2.0 / np.sqrt(d))
  pi = rng.dirichlet(np.ones(k))
  comp = rng.choice(k, size=n, p=pi)
  return means[comp] + rng.normal(size=(n, d)) / np.sqrt(d)


def gmm_pair(seed: int, n: int = 2048, m: int = 2048, d: int = 16, k: int = 4,
             eps_mult: float = 0.05) -> Problem:
  """x and y: k-component mixtures drawn in turn from one generator (seed); uniform weights.

  The learned warm start's d=16 family: its training problems use seeds 1000..20999
  (with eps_mult drawn per problem), held-out evaluation seeds start at 100000.
  """
  rng = np.random.default_rng(seed)
  x = gaussian_mixture(rng, n, d, k)
  y = gaussian_mixture(rng, m, d, k)
  return Problem(x.astype(np.float32), y.astype(np.float32), None, None, eps_mult,
                 name=f"gmm_pair_seed{seed}_n{n}_m{m}_d{d}")


def mixture_pair(n: int = 32768, m: int = 8192, d: int = 128, k: int = 8,
                 seed_x: int = 2, seed_y: int = 3, tau: float = 0.95,
                 eps_mult: float = 0.02, weights: str = "uniform_noise") -> Problem:
  """Two separately seeded k-component mixtures; a, b = U(0.5, 1.5) normalised."""
  rx = np.random.default_rng(seed_x)
  ry = np.random.default_rng(seed_y)
  x = gaussian_mixture(rx, n, d, k)
  y = gaussian_mixture(ry, m, d, k)
  if weights == "uniform":
    a = b = None
  else:
    a = _normalise(rx.uniform(0.5, 1.5, size=n))
    b = _normalise(ry.uniform(0.5, 1.5, size=m))
  return Problem(x.astype(np.float32), y.astype(np.float32), a, b, eps_mult, tau, tau,
                 name=f"mixture_pair_n{n}_m{m}_d{d}_k{k}")


def batched_uneven(count: int = 512, d: int = 32, lo: int = 64, hi: int = 4096,
                   seed: int = 4) -> List[Problem]:
  """Problems with n_k, m_k log-uniform in [lo, hi]; y_k shifted by N(0, 0.25 I/d)."""
  rng = np.random.default_rng(seed)
  out = []
  for i in range(count):
    n_k = int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    m_k = int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    delta = rng.normal(size=(d,)) * (0.5 / np.sqrt(d))
    x = rng.normal(size=(n_k, d)) / np.sqrt(d)
    y = rng.normal(size=(m_k, d)) / np.sqrt(d) + delta
    out.append(Problem(x.astype(np.float32), y.astype(np.float32), None, None, 0.05,
                       name=f"batched_{i}"))
  return out
