"""FP64 CPU oracle: OTT's log-domain Sinkhorn on a squared-Euclidean point cloud.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: it may
be imported only by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``.  Nothing in ``ott_amd`` imports it, and
the product path never routes through it.

Every function restates, operation for operation, the reference function it
cites (paths are relative to the ott-jax/ott source tree ``src/ott``).  The
arithmetic is float64 by default; passing ``dtype=np.float32`` gives an
op-order twin that keeps OTT's float32 operation order ((f + g - C) / eps, then
a max-shifted log-sum-exp, expanded squared-Euclidean cost), so float32-level
agreement with the GPU path can be shown as well.

Pinning (what checks this oracle independently of its own formulas) lives in
``tests/test_oracle.py``: brute-force costs, ``scipy.special.logsumexp``, the
m=1 closed form of a half-step, the exact marginal after a half-step, an
L-BFGS solve of the smooth semi-dual, the Hungarian limit at small epsilon, the
epsilon -> infinity limit, scale invariance and gauge invariance; and in
``tests/test_oracle_costs.py`` for the unbalanced case: the fixed point against
an L-BFGS minimisation of its primal, ``compute_kl_reg_cost`` against a
numerical infimum of its Lagrangian, closed forms for one point per side,
continuity at tau -> 1, planted slips that must fail, and ``ent_reg_cost`` /
``dual_cost`` against the coupling's own entropy; for recentering (both taus
< 1), ``recenter`` mapping any translate of the optimum back to it, each
recentred half-step returning the translation-optimal potential (an identity
that uses neither k, xi nor smin), the recentred iteration reaching the plain
iteration's optimum, and planted slips in k, xi and smin.

One property of the reference is kept as it is: with tau < 1 its iteration
converges to the optimum of the problem whose entropy is taken against the
counting measure, while ``compute_kl_reg_cost`` is the dual function of the
problem whose entropy is taken against a b^T, evaluated at f - eps log a and
g - eps log b.  So in the unbalanced case reg_ot_cost is not the primal value
of the returned coupling (it is in the balanced case), and this oracle, like
the HIP path, reproduces the reference's number rather than that primal.
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Optional, Sequence, Tuple

import numpy as np
from scipy import special as sps

__all__ = [
    "sqeuclidean_cost",
    "ti_cost",
    "mean_cost",
    "std_cost",
    "median_cost",
    "resolve_epsilon",
    "epsilon_at",
    "momentum_weight",
    "reg_ot_cost_grad",
    "DenseLSE",
    "apply_lse_kernel",
    "update_potential",
    "marginal_from_potentials",
    "grad_of_marginal_fit",
    "marginal_error",
    "solution_error",
    "compute_kl_reg_cost",
    "recenter",
    "recenter_k",
    "recenter_xi",
    "smin",
    "OracleOutput",
    "sinkhorn",
    "ent_reg_cost",
    "dual_cost",
    "segment_point_cloud",
    "segment_sinkhorn",
    "transport_from_potentials",
    "apply_transport_from_potentials",
    "transport_cost_at_cost",
    "soft_c_transform",
    "barycentric_map",
]

DEFAULT_EPSILON_SCALE = 0.05  # geometry/epsilon_scheduler.py:22


# --------------------------------------------------------------------------
# Costs and epsilon
# --------------------------------------------------------------------------
def sqeuclidean_cost(x, y, scale: float = 1.0, *, expanded: bool = False,
                     dtype=np.float64) -> np.ndarray:
  """Squared-Euclidean cost matrix ``s * ||x_i - y_j||^2``.

  geometry/costs.py:554-561 (``norm(x) + norm(y) - 2 vdot``) applied pairwise
  (costs.py:82-92) and scaled by ``inv_scale_cost`` (pointcloud.py:334-340).
  ``expanded=True`` keeps the reference's expanded form; the default forms the
  differences directly, which is exact up to one rounding per term.
  """
  x = np.asarray(x, dtype=dtype)
  y = np.asarray(y, dtype=dtype)
  if expanded:
    nx = np.sum(x ** 2, axis=-1)
    ny = np.sum(y ** 2, axis=-1)
    c = nx[:, None] + ny[None, :] + (-2.0 * (x @ y.T)).astype(dtype)
  else:
    diff = x[:, None, :] - y[None, :, :]
    c = np.sum(diff * diff, axis=-1)
  return (np.asarray(scale, dtype=dtype) * c).astype(dtype)


def ti_cost(x, y, cost_fn, scale: float = 1.0) -> np.ndarray:
  """Cost matrix ``s * c(x_i, y_j)`` of a named cost (geometry/costs.py):
  ("sqeuclidean",) ``||z||^2`` (:548-573), ("pnormp", p) ``||z||_p^p / p``
  (:304-334), ("sqpnorm", p) ``||z||_p^2 / 2`` (:269-301), ("euclideanp", p)
  ``||z||_2^p`` (:337-361), ("euclidean",) ``||z||_2`` (:528-545), with
  z = x_i - y_j (TICost.__call__, :155-157) and the p-norm of ``mu.norm``."""
  x = np.asarray(x, dtype=np.float64)
  y = np.asarray(y, dtype=np.float64)
  kind = cost_fn[0]
  z = np.abs(x[:, None, :] - y[None, :, :])
  if kind == "sqeuclidean":
    c = np.sum(z * z, axis=-1)
  elif kind == "euclidean":
    c = np.sqrt(np.sum(z * z, axis=-1))
  else:
    p = float(cost_fn[1])
    if kind == "pnormp":
      c = np.sum(z ** p, axis=-1) / p
    elif kind == "sqpnorm":
      c = 0.5 * np.sum(z ** p, axis=-1) ** (2.0 / p)
    elif kind == "euclideanp":
      c = np.sqrt(np.sum(z * z, axis=-1)) ** p
    else:
      raise ValueError(f"unknown cost {cost_fn!r}")
  return scale * c


def mean_cost(x, y, scale: float = 1.0) -> float:
  """Mean of the cost matrix in closed form.

  geometry/geometry.py:112-117 calls ``apply_cost`` with a uniform vector, which
  PointCloud answers with pointcloud.py:170-186:
  ``sum_j (1/m) * s * (nx . vec + ny * sum(vec) - 2 y . (x^T vec))``.
  """
  x = np.asarray(x, dtype=np.float64)
  y = np.asarray(y, dtype=np.float64)
  n, m = x.shape[0], y.shape[0]
  vec = np.full((n,), 1.0 / n)
  nx = np.sum(x ** 2, axis=-1)
  ny = np.sum(y ** 2, axis=-1)
  applied = nx @ vec + ny * np.sum(vec) - 2.0 * (y @ (x.T @ vec))
  return float(np.sum((1.0 / m) * scale * applied))


def std_cost(x, y, scale: float = 1.0) -> float:
  """Standard deviation of the cost matrix (geometry/geometry.py:119-136)."""
  c = sqeuclidean_cost(x, y, scale)
  n, m = c.shape
  tmp = np.sum((1.0 / m) * np.sum(c ** 2, axis=0) / n) - mean_cost(x, y, scale) ** 2
  return float(np.sqrt(max(tmp, 0.0)))


def median_cost(x, y, scale: float = 1.0) -> float:
  """Median of the cost matrix (geometry/geometry.py:106-109, jnp.median:
  the mean of the two middle entries when n m is even), the value that
  scale_cost="median" divides by (geometry/pointcloud.py:351-353)."""
  return float(np.median(sqeuclidean_cost(x, y, scale)))


def resolve_epsilon(x, y, epsilon=None, relative_epsilon=None,
                    scale: float = 1.0) -> float:
  """Target epsilon of a PointCloud (geometry/geometry.py:147-176)."""
  if relative_epsilon is None:
    if epsilon is not None:
      return float(epsilon)
    return DEFAULT_EPSILON_SCALE * std_cost(x, y, scale)
  if relative_epsilon == "std":
    stat = std_cost(x, y, scale)
  elif relative_epsilon == "mean":
    stat = mean_cost(x, y, scale)
  else:
    raise ValueError(f"Invalid relative epsilon: {relative_epsilon}.")
  mult = DEFAULT_EPSILON_SCALE if epsilon is None else epsilon
  return float(mult * stat)


def epsilon_at(target: float, it: Optional[int], init: float = 1.0,
               decay: float = 1.0) -> float:
  """Scheduled epsilon (geometry/epsilon_scheduler.py:47-60)."""
  if it is None:
    return target
  return max(init * decay ** it, 1.0) * target


# --------------------------------------------------------------------------
# Log-domain kernel operators (the Geometry protocol)
# --------------------------------------------------------------------------
def _lse(z: np.ndarray, axis: int) -> np.ndarray:
  """Max-shifted log-sum-exp; -inf for an all -inf slice.

  math/utils.py:132-138 forwards to ``jax.scipy.special.logsumexp``; scipy's
  routine computes the same max-shifted quantity.
  """
  with np.errstate(invalid="ignore", over="ignore"):
    return sps.logsumexp(z, axis=axis)


class DenseLSE:
  """``C``-backed implementation of the Geometry LSE protocol.

  ``lse(f, g, eps, axis)`` returns ``eps * LSE((f[:, None] + g[None, :] - C) / eps)``
  over ``axis`` (geometry/geometry.py:423-442, ``_center`` then ``_softmax``).
  """

  def __init__(self, cost: np.ndarray):
    self.cost = cost

  @property
  def shape(self) -> Tuple[int, int]:
    return self.cost.shape

  def lse(self, f, g, eps, axis):
    dt = self.cost.dtype
    with np.errstate(invalid="ignore"):
      centered = f.astype(dt)[:, None] + g.astype(dt)[None, :] - self.cost
      return (np.asarray(eps, dt) * _lse(centered / np.asarray(eps, dt), axis)
              ).astype(dt)


class PointCloudLSE:
  """Cost-free implementation of the same protocol for large clouds.

  Backed by ``liboracle_sinkhorn.so`` (oracle/sinkhorn_cpu.c, FP64, OpenMP):
  the online mode of pointcloud.py:85-103 with the cost recomputed per block.
  """

  def __init__(self, x, y, scale: float = 1.0, threads: int = 0):
    import ctypes
    import os
    lib_path = os.path.join(os.path.dirname(__file__), "liboracle_sinkhorn.so")
    if not os.path.exists(lib_path):
      raise RuntimeError(f"oracle C library not built: {lib_path} (run make -C oracle)")
    self._lib = ctypes.CDLL(lib_path)
    self._lib.ox_lse.restype = ctypes.c_int
    self._ct = ctypes
    self.x = np.ascontiguousarray(x, dtype=np.float64)
    self.y = np.ascontiguousarray(y, dtype=np.float64)
    self.scale = float(scale)
    self.threads = int(threads)

  @property
  def shape(self) -> Tuple[int, int]:
    return self.x.shape[0], self.y.shape[0]

  def lse(self, f, g, eps, axis):
    ct = self._ct
    n, m = self.shape
    f = np.ascontiguousarray(f, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty((m if axis == 0 else n,), dtype=np.float64)
    dp = ct.POINTER(ct.c_double)
    rc = self._lib.ox_lse(
        self.x.ctypes.data_as(dp), self.y.ctypes.data_as(dp),
        ct.c_int64(n), ct.c_int64(m), ct.c_int32(self.x.shape[1]),
        ct.c_double(self.scale), f.ctypes.data_as(dp), g.ctypes.data_as(dp),
        ct.c_double(float(eps)), ct.c_int(axis), out.ctypes.data_as(dp),
        ct.c_int(self.threads))
    if rc != 0:
      raise ValueError("ox_lse rejected its arguments")
    return out


def apply_lse_kernel(ops, f, g, eps, axis=0):
  """geometry/geometry.py:260-299 (vec=None): softmax minus the finite h."""
  res = ops.lse(f, g, eps, axis)
  remove = f if axis == 1 else g
  with np.errstate(invalid="ignore"):
    return res - np.where(np.isfinite(remove), remove, 0.0)


def update_potential(ops, f, g, log_marginal, eps, axis=0):
  """geometry/geometry.py:376-398."""
  app = apply_lse_kernel(ops, f, g, eps, axis)
  with np.errstate(invalid="ignore"):
    return eps * log_marginal - np.where(np.isfinite(app), app, 0.0)


def marginal_from_potentials(ops, f, g, eps, axis=0):
  """geometry/geometry.py:326-349 (always the target epsilon)."""
  h = f if axis == 1 else g
  z = apply_lse_kernel(ops, f, g, eps, axis)
  with np.errstate(invalid="ignore", over="ignore"):
    return np.exp((z + h) / eps)


# --------------------------------------------------------------------------
# Errors and costs
# --------------------------------------------------------------------------
def rho(eps, tau):
  """math/unbalanced_functions.py:89-90."""
  return (eps * tau) / (1.0 - tau)


def phi_star(h, r):
  """math/unbalanced_functions.py:19-21."""
  return r * (np.exp(h / r) - 1.0)


def grad_of_marginal_fit(c, h, tau, eps):
  """math/unbalanced_functions.py:30-50."""
  if tau == 1.0:
    return c
  r = rho(eps, tau)
  with np.errstate(over="ignore", invalid="ignore"):
    return np.where(c > 0, c * np.exp(-h / r), 0.0)


def marginal_error(ops, f, g, target, eps, axis=0, norm_error=(1,)):
  """solvers/linear/sinkhorn.py:191-221."""
  marginal = marginal_from_potentials(ops, f, g, eps, axis)
  p = np.asarray(norm_error, dtype=marginal.dtype)
  return np.sum(np.abs(marginal - target)[None, :] ** p[:, None], axis=1) ** (1.0 / p)


def solution_error(ops, f, g, a, b, eps, tau_a, tau_b, norm_error=(1,),
                   parallel_dual_updates=False):
  """solvers/linear/sinkhorn.py:127-188 (lse_mode=True)."""
  if tau_a == 1.0 and tau_b == 1.0 and not parallel_dual_updates:
    return marginal_error(ops, f, g, b, eps, 0, norm_error)
  grad_a = grad_of_marginal_fit(a, f, tau_a, eps)
  grad_b = grad_of_marginal_fit(b, g, tau_b, eps)
  err = marginal_error(ops, f, g, grad_a, eps, 1, norm_error)
  err = err + marginal_error(ops, f, g, grad_b, eps, 0, norm_error)
  return err


def compute_kl_reg_cost(ops, f, g, a, b, eps, tau_a=1.0, tau_b=1.0):
  """solvers/linear/sinkhorn.py:224-277 (lse_mode=True).

  For tau < 1 this is the Lagrange dual function of the a b^T-referenced
  unbalanced problem at (f - eps log a, g - eps log b), whatever (f, g) is;
  at the solver's fixed point it differs from the primal value of the
  coupling (see the module docstring)."""
  supp_a, supp_b = a > 0, b > 0
  with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
    fa = eps * np.log(a)  # geometry.py:524-534 potential_from_scaling
    if tau_a == 1.0:
      div_a = np.sum(np.where(supp_a, a * (f - fa), 0.0))
    else:
      div_a = -np.sum(np.where(supp_a, a * phi_star(-(f - fa), rho(eps, tau_a)), 0.0))
    gb = eps * np.log(b)
    if tau_b == 1.0:
      div_b = np.sum(np.where(supp_b, b * (g - gb), 0.0))
    else:
      div_b = -np.sum(np.where(supp_b, b * phi_star(-(g - gb), rho(eps, tau_b)), 0.0))
  total_sum = np.sum(marginal_from_potentials(ops, f, g, eps, 0))
  return div_a + div_b + eps * (np.sum(a) * np.sum(b) - total_sum)


def ent_reg_cost(reg_ot_cost, a, b, eps):
  """solvers/linear/sinkhorn.py:350-370: reg cost minus eps * (H(a) + H(b))."""
  return reg_ot_cost - eps * (np.sum(sps.entr(a)) + np.sum(sps.entr(b)))


def dual_cost(f, g, a, b):
  """solvers/linear/sinkhorn.py:337-343."""
  return (np.sum(np.where(a > 0.0, a * f, 0.0)) +
          np.sum(np.where(b > 0.0, b * g, 0.0)))


def recenter(f, g, a, b, eps, tau_a, tau_b):
  """solvers/linear/sinkhorn.py:76-114."""
  if tau_a == 1.0 and tau_b == 1.0:
    fin = np.isfinite(f)
    shift = np.sum(np.where(fin, f, 0.0)) / np.sum(fin)
    return f - shift, g + shift
  if tau_a == 1.0 or tau_b == 1.0:
    return f, g
  ra, rb = rho(eps, tau_a), rho(eps, tau_b)
  tau = ra * rb / (ra + rb)
  shift = tau * (sps.logsumexp(-f / ra, b=a) - sps.logsumexp(-g / rb, b=b))
  return f + shift, g - shift


def recenter_k(tau_i, tau_j, tau_a, tau_b):
  """``k`` of Sinkhorn.lse_step (sinkhorn.py:809-812)."""
  num = -tau_j * (tau_a - 1.0) * (tau_b - 1.0) * (tau_i - 1.0)
  denom = (tau_j - 1.0) * (tau_a * (tau_b - 1.0) + tau_b * (tau_a - 1.0))
  return num / denom


def recenter_xi(tau_i, tau_j, tau_a, tau_b):
  """``xi`` of Sinkhorn.lse_step (sinkhorn.py:814-816)."""
  k_ij = recenter_k(tau_i, tau_j, tau_a, tau_b)
  return k_ij / (1.0 - k_ij)


def smin(potential, marginal, eps, tau):
  """``smin`` of Sinkhorn.lse_step (sinkhorn.py:818-822): the soft minimum
  -rho log sum_i marginal_i exp(-potential_i / rho)."""
  r = rho(eps, tau)
  return -r * sps.logsumexp(-potential / r, b=marginal)


# --------------------------------------------------------------------------
# The solver loop
# --------------------------------------------------------------------------
@dataclasses.dataclass
class OracleOutput:
  f: np.ndarray
  g: np.ndarray
  errors: np.ndarray
  reg_ot_cost: float
  converged: bool
  n_iters: int
  epsilon: float
  threshold: float
  inner_iterations: int


def momentum_weight(errors, it, value, start, error_threshold, inner_iterations):
  """Momentum.weight and Momentum.lehmann (acceleration.py:138-160): ``value``
  before iteration ``start`` (or when start = 0), then, if the error of block
  start // inner - 1 is below ``error_threshold``, Lehmann's weight
  2 / (1 + sqrt(1 - r^(1 / inner))) with r = min(err[idx-1] / err[idx-2], 0.99)
  in float32, as the reference evaluates it on its float32 errors."""
  if start == 0 or it < start:
    return value
  idx = start // inner_iterations
  e1 = np.float32(errors[idx - 1])
  if not e1 < error_threshold:
    return value
  ratio = np.minimum(e1 / np.float32(errors[idx - 2]), np.float32(0.99))
  power = np.float32(1.0 / inner_iterations)
  return float(np.float32(2.0) / (np.float32(1.0) + np.sqrt(np.float32(1.0) - ratio ** power)))


def sinkhorn(
    ops,
    a: np.ndarray,
    b: np.ndarray,
    epsilon: float,
    *,
    tau_a: float = 1.0,
    tau_b: float = 1.0,
    threshold: float = 1e-3,
    norm_error: int = 1,
    inner_iterations: int = 10,
    min_iterations: int = 0,
    max_iterations: int = 2000,
    momentum: float = 1.0,
    parallel_dual_updates: bool = False,
    recenter_potentials: bool = False,
    init: Optional[Tuple[np.ndarray, np.ndarray]] = None,
    eps_init: float = 1.0,
    eps_decay: float = 1.0,
    momentum_start: int = 0,
    momentum_error_threshold: float = np.inf,
    progress_fn: Optional[Callable] = None,
    dtype=np.float64,
) -> OracleOutput:
  """Log-domain Sinkhorn restated from OTT.

  - initializer: initializers/linear/initializers.py:70-131 (zeros, then -inf
    on zero weights) unless ``init`` is given (sinkhorn.py:782-801);
  - loop: math/fixed_point_loop.py:23-93 (blocks of ``inner_iterations``,
    stop test between blocks, ``lax.scan`` when min == max);
  - one iteration: sinkhorn.py:879-938 with ``lse_step`` sinkhorn.py:803-857
    (with ``recenter_potentials`` and both taus < 1, the translation-invariant
    terms -k smin(other) + xi smin(new) of sinkhorn.py:824-854, rho at the
    target epsilon as ``ot_prob.epsilon`` is) and ``Momentum``, constant or
    adaptive (acceleration.py:138-172);
  - termination: sinkhorn.py:940-962; outputs: sinkhorn.py:974-1029, n_iters
    sinkhorn.py:431-434; cost: sinkhorn.py:224-277 through ``run``
    sinkhorn.py:1052-1062.
  """
  rc_terms = recenter_potentials and tau_a < 1.0 and tau_b < 1.0
  a = np.asarray(a, dtype=dtype)
  b = np.asarray(b, dtype=dtype)
  eps_t = np.asarray(epsilon, dtype=dtype)
  n, m = ops.shape
  with np.errstate(divide="ignore"):
    log_a, log_b = np.log(a), np.log(b)
  if init is None:
    f = np.where(a > 0.0, 0.0, -np.inf).astype(dtype)
    g = np.where(b > 0.0, 0.0, -np.inf).astype(dtype)
  else:
    f = np.asarray(init[0], dtype=dtype).copy()
    g = np.asarray(init[1], dtype=dtype).copy()

  outer = int(np.ceil(max_iterations / inner_iterations))
  errors = -np.ones((outer,), dtype=dtype)
  force_scan = min_iterations == max_iterations

  def one_iteration(it, f, g, compute_error):
    eps_it = dtype(epsilon_at(float(epsilon), it, eps_init, eps_decay))
    w = dtype(momentum_weight(errors, it, momentum, momentum_start, momentum_error_threshold,
                              inner_iterations))
    old_f, old_g = f, g
    new_g = tau_b * update_potential(ops, old_f, old_g, log_b, eps_it, axis=0)
    if rc_terms:
      new_g = new_g - recenter_k(tau_b, tau_b, tau_a, tau_b) * smin(old_f, a, eps_t, tau_a)
      new_g = new_g + recenter_xi(tau_b, tau_a, tau_a, tau_b) * smin(new_g, b, eps_t, tau_b)
    with np.errstate(invalid="ignore"):
      g = (1.0 - w) * np.where(np.isfinite(old_g), old_g, 0.0) + w * new_g
    g_for_f = old_g if parallel_dual_updates else g
    new_f = tau_a * update_potential(ops, old_f, g_for_f, log_a, eps_it, axis=1)
    if rc_terms:
      new_f = new_f - recenter_k(tau_a, tau_a, tau_a, tau_b) * smin(g_for_f, b, eps_t, tau_b)
      new_f = new_f + recenter_xi(tau_a, tau_b, tau_a, tau_b) * smin(new_f, a, eps_t, tau_a)
    with np.errstate(invalid="ignore"):
      f = (1.0 - w) * np.where(np.isfinite(old_f), old_f, 0.0) + w * new_f
    if it == max_iterations - 1 or (compute_error and it >= min_iterations):
      ff, gg = f, g
      if recenter_potentials:
        ff, gg = recenter(f, g, a, b, eps_t, tau_a, tau_b)
      err = solution_error(ops, ff, gg, a, b, eps_t, tau_a, tau_b,
                           (norm_error,), parallel_dual_updates)[0]
    else:
      err = np.inf
    errors[it // inner_iterations] = err
    if progress_fn is not None:
      progress_fn((it, inner_iterations, max_iterations, errors))
    return f, g

  def run_block(it, f, g):
    for j in range(inner_iterations):
      f, g = one_iteration(it, f, g, j == inner_iterations - 1)
      it += 1
    return it, f, g

  it = 0
  if force_scan:
    for _ in range(max_iterations // inner_iterations):
      it, f, g = run_block(it, f, g)
  else:
    while True:
      if not it < max_iterations:
        break
      if not it < min_iterations:
        err = errors[it // inner_iterations - 1]
        converged = it > 0 and err < threshold
        diverged = not np.isfinite(err)
        if converged or diverged:
          break
      it, f, g = run_block(it, f, g)

  if recenter_potentials:
    f, g = recenter(f, g, a, b, eps_t, tau_a, tau_b)
  converged = bool((not np.any(np.isnan(errors))) and errors[-1] < threshold)
  n_iters = int(np.sum(errors != -1) * inner_iterations)
  reg = compute_kl_reg_cost(ops, f, g, a, b, eps_t, tau_a, tau_b)
  return OracleOutput(f=f, g=g, errors=errors, reg_ot_cost=float(reg),
                      converged=converged, n_iters=n_iters,
                      epsilon=float(epsilon), threshold=threshold,
                      inner_iterations=inner_iterations)


def solve_pointcloud(x, y, a=None, b=None, *, epsilon=None,
                     relative_epsilon=None, scale_cost: float = 1.0,
                     dtype=np.float64, expanded: bool = False, **kwargs):
  """``linear.solve(PointCloud(x, y, ...), a, b, **kwargs)`` with a dense cost.

  solvers/linear/_solve.py:29-64 and problems/linear/linear_problem.py:65-79
  (``None`` weights become uniform).
  """
  x = np.asarray(x, dtype=np.float64)
  y = np.asarray(y, dtype=np.float64)
  n, m = x.shape[0], y.shape[0]
  a = np.full((n,), 1.0 / n) if a is None else np.asarray(a, np.float64)
  b = np.full((m,), 1.0 / m) if b is None else np.asarray(b, np.float64)
  cost_fn = kwargs.pop("cost_fn", None)
  if cost_fn is None:
    eps = resolve_epsilon(x, y, epsilon, relative_epsilon, 1.0 / scale_cost)
    cost = sqeuclidean_cost(x, y, 1.0 / scale_cost, expanded=expanded, dtype=dtype)
  else:
    # geometry.py:147-171 on the dense matrix of another cost
    cost = ti_cost(x, y, cost_fn, 1.0 / scale_cost)
    stat = {"mean": np.mean(cost), "std": np.std(cost), None: np.std(cost)}[relative_epsilon]
    if relative_epsilon is None and epsilon is not None:
      eps = float(epsilon)
    else:
      eps = float((DEFAULT_EPSILON_SCALE if epsilon is None else epsilon) * stat)
    cost = cost.astype(dtype)
  return sinkhorn(DenseLSE(cost), a, b, eps, dtype=dtype, **kwargs)


def segment_point_cloud(x, a, num_per_segment, max_measure_size, padding_vector=None):
  """geometry/segment.py:23-124, second interface (contiguous segments).

  Returns padded points [K, P, d] and weights [K, P] with P =
  min(max_measure_size, num); rows beyond P in a segment are dropped, as in
  the reference's dynamic_slice.
  """
  x = np.asarray(x, dtype=np.float64)
  num, dim = x.shape
  counts = np.asarray(num_per_segment, dtype=np.int64)
  P = min(int(max_measure_size), num)
  if a is None:
    a = np.repeat(1.0 / counts, counts)
  a = np.asarray(a, dtype=np.float64)
  pad = np.zeros((dim,)) if padding_vector is None else np.asarray(padding_vector, np.float64).reshape(dim)
  seg_x = np.tile(pad, (len(counts), P, 1))
  seg_a = np.zeros((len(counts), P))
  off = 0
  for k, c in enumerate(counts):
    take = min(int(c), P)
    seg_x[k, :take] = x[off:off + take]
    seg_a[k, :take] = a[off:off + take]
    off += int(c)
  return seg_x, seg_a


def segment_sinkhorn(x, y, num_per_segment_x, num_per_segment_y, max_measure_size=None,
                     weights_x=None, weights_y=None, sinkhorn_kwargs=None, **pc_kwargs):
  """tools/segment_sinkhorn.py:24-138: pad every segment, solve each padded pair.

  Returns one OracleOutput per segment, on the padded problem (padding rows
  carry zero weight and -inf potentials).
  """
  if max_measure_size is None:
    max_measure_size = max(max(num_per_segment_x), max(num_per_segment_y))
  sx, sa = segment_point_cloud(x, weights_x, num_per_segment_x, max_measure_size)
  sy, sb = segment_point_cloud(y, weights_y, num_per_segment_y, max_measure_size)
  outs = []
  for k in range(sx.shape[0]):
    outs.append(solve_pointcloud(sx[k], sy[k], sa[k], sb[k], **pc_kwargs, **(sinkhorn_kwargs or {})))
  return outs


# --------------------------------------------------------------------------
# Output side: the transport matrix and its applications
# --------------------------------------------------------------------------

def transport_from_potentials(cost: np.ndarray, f, g, eps: float) -> np.ndarray:
  """geometry.py:361-365 with _center (geometry.py:351-352): exp((f_i + g_j - C_ij) / eps)."""
  f = np.asarray(f, dtype=np.float64)
  g = np.asarray(g, dtype=np.float64)
  with np.errstate(invalid="ignore"):
    return np.exp((f[:, None] + g[None, :] - cost) / eps)


def apply_transport_from_potentials(cost: np.ndarray, f, g, vec, eps: float, axis: int = 0) -> np.ndarray:
  """geometry.py:445-486 (lse_mode): axis 0 sums over the rows, P^T vec; axis 1
  over the columns, P vec; vec is one vector or a batch [k, .].  The reference
  takes a signed, weighted log-sum-exp; this restatement forms the sum
  directly in float64."""
  p = transport_from_potentials(cost, f, g, eps)
  v = np.asarray(vec, dtype=np.float64)
  return (v @ p) if axis == 0 else (v @ p.T)


def soft_c_transform(cost: np.ndarray, fg, weights, eps: float, axis: int) -> np.ndarray:
  """LinearProblem._c_transform, soft branch (linear_problem.py:149-171):
  fg expanded on axis 1 - axis, z = (fg - C) / eps, -eps logsumexp(z, b=weights,
  axis=axis) with ``weights`` broadcast against z as the reference's does."""
  fg = np.expand_dims(np.asarray(fg, np.float64), 1 - axis)
  z = (fg - cost) / eps
  w = np.broadcast_to(np.asarray(weights, np.float64), z.shape)
  return -eps * sps.logsumexp(z, b=w, axis=axis)


def barycentric_map(cost: np.ndarray, h, weights, eps: float, support) -> np.ndarray:
  """x - grad f(x) / 2 for f(x) = -eps log sum_k w_k exp((h_k - |x - z_k|^2) / eps)
  (potentials.py:52-84, costs.py:220-226): the pi-weighted mean of the support,
  pi_k(x) proportional to w_k exp((h_k - C(x, z_k)) / eps); ``cost`` is [k_new, K]."""
  z = (np.asarray(h, np.float64)[None, :] - cost) / eps
  with np.errstate(divide="ignore"):
    z = z + np.log(np.broadcast_to(np.asarray(weights, np.float64), z.shape))
  pi = np.exp(z - sps.logsumexp(z, axis=1, keepdims=True))
  return pi @ np.asarray(support, np.float64)


def transport_cost_at_cost(cost_p: np.ndarray, f, g, eps: float, other_cost: np.ndarray) -> float:
  """sinkhorn.py:386-411 (dense branch): sum(P * C'), P from the solve's cost."""
  return float((transport_from_potentials(cost_p, f, g, eps) * other_cost).sum())



def reg_ot_cost_grad(x, y, a, b, f, g, eps, scale=1.0):
  """Gradient of reg_ot_cost with f and g held fixed (the Danskin gradient of
  sinkhorn.py:329-335 through compute_kl_reg_cost, sinkhorn.py:224-277):
  with C = scale |x_i - y_j|^2 and P = exp((f + g - C) / eps),
    d/dx_i = 2 scale sum_j P_ij (x_i - y_j),  d/dy_j = 2 scale sum_i P_ij (y_j - x_i),
    d/da_i = f_i - eps log a_i - eps + eps sum(b)  (a_i > 0; 0 otherwise),
  and d/db likewise.  Balanced problems, eps and scale held constant.

  At a zero weight a_i = 0 the a-gradient here is eps sum(b) (the where()
  branch), a value this library chose: jax.grad through the reference's
  compute_kl_reg_cost gives NaN there (a (f - eps log a) with f = -inf and
  1 / a from log a pass a NaN through jnp.where), so parity at zero weights is
  a documented deviation, not a pinned value."""
  x = np.asarray(x, np.float64)
  y = np.asarray(y, np.float64)
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  p = transport_from_potentials(sqeuclidean_cost(x, y, scale), f, g, eps)
  gx = 2.0 * scale * (p.sum(1)[:, None] * x - p @ y)
  gy = 2.0 * scale * (p.sum(0)[:, None] * y - p.T @ x)
  with np.errstate(divide="ignore", invalid="ignore"):
    ga = np.where(a > 0, f - eps * np.log(a) - eps, 0.0) + eps * b.sum()
    gb = np.where(b > 0, g - eps * np.log(b) - eps, 0.0) + eps * a.sum()
  return gx, gy, ga, gb
