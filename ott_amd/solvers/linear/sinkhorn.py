"""Log-domain Sinkhorn solver on the ottx HIP kernels.

Mirrors ``ott.solvers.linear.sinkhorn`` (solvers/linear/sinkhorn.py): the
``Sinkhorn`` constructor (:718-780), ``__call__`` (:782-801), and
``SinkhornOutput`` (:280-538).  The whole loop -- fixed-point blocks
(math/fixed_point_loop.py:23-93), lse steps (:803-857), errors (:127-221),
termination (:940-962), outputs (:974-1029) and ``reg_ot_cost`` (:224-277) --
runs in ``ottx_sinkhorn_solve`` (include/ottx.h); this module validates the
options, prepares buffers and wraps the results.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any, Callable, NamedTuple, Optional, Tuple

import numpy as np
import torch

from ott_amd import _lib, tune
from ott_amd.geometry.pointcloud import PointCloud, as_device_f32
from ott_amd.initializers.linear import initializers as init_lib
from ott_amd.problems.linear import linear_problem
from ott_amd.solvers.linear import acceleration

__all__ = ["Sinkhorn", "SinkhornOutput"]


class SinkhornOutput(NamedTuple):
  """sinkhorn.py:280-323: potentials, errors, reg_ot_cost, ot_prob, threshold,
  converged, inner_iterations."""

  potentials: Tuple[torch.Tensor, torch.Tensor]
  errors: Optional[torch.Tensor] = None
  reg_ot_cost: Optional[float] = None
  ot_prob: Optional[linear_problem.LinearProblem] = None
  threshold: Optional[float] = None
  converged: Optional[bool] = None
  inner_iterations: Optional[int] = None
  # the count the solve itself reported (beyond OTT's fields): n_iters without
  # a reduction over ``errors`` and its device-to-host wait
  solve_n_iters: Optional[int] = None

  def set(self, **kwargs: Any) -> "SinkhornOutput":
    if "errors" in kwargs or "inner_iterations" in kwargs:
      kwargs.setdefault("solve_n_iters", None)  # no longer the count of these errors
    return self._replace(**kwargs)

  @property
  def f(self) -> torch.Tensor:
    return self.potentials[0]

  @property
  def g(self) -> torch.Tensor:
    return self.potentials[1]

  @property
  def geom(self) -> PointCloud:
    return self.ot_prob.geom

  @property
  def a(self) -> torch.Tensor:
    return self.ot_prob.a

  @property
  def b(self) -> torch.Tensor:
    return self.ot_prob.b

  @property
  def n_iters(self) -> int:
    """sinkhorn.py:431-434."""
    if self.solve_n_iters is not None:
      return self.solve_n_iters
    return int(torch.sum(self.errors != -1).item()) * self.inner_iterations

  @property
  def dual_cost(self) -> float:
    """sinkhorn.py:337-343."""
    a, b = self.ot_prob.a, self.ot_prob.b
    fa = torch.where(a > 0.0, a.double() * self.f.double(), torch.zeros_like(a, dtype=torch.float64))
    gb = torch.where(b > 0.0, b.double() * self.g.double(), torch.zeros_like(b, dtype=torch.float64))
    return float(fa.sum() + gb.sum())

  @property
  def ent_reg_cost(self) -> float:
    """sinkhorn.py:350-370: ``reg_ot_cost - eps * (H(a) + H(b))``."""
    def entr(w):
      w = w.double()
      return torch.where(w > 0, -w * torch.log(w), torch.zeros_like(w)).sum()
    return float(self.reg_ot_cost - self.geom.epsilon * (entr(self.a) + entr(self.b)))

  @property
  def kl_reg_cost(self) -> float:
    return self.reg_ot_cost

  def marginal(self, axis: int) -> torch.Tensor:
    """sinkhorn.py:451-452 -> geometry.marginal_from_potentials."""
    return self.ot_prob.geom.marginal_from_potentials(self.f, self.g, axis=axis)

  @property
  def transport_mass(self) -> float:
    return float(self.marginal(0).double().sum())

  @property
  def scalings(self) -> Tuple[torch.Tensor, torch.Tensor]:
    """sinkhorn.py:437-441: exp(f / eps), exp(g / eps)."""
    eps = self.geom.epsilon
    return torch.exp(self.f / eps), torch.exp(self.g / eps)

  @property
  def matrix(self) -> torch.Tensor:
    """sinkhorn.py:443-449: the dense transport matrix (geometry.py:361-365)."""
    return self.geom.transport_from_potentials(self.f, self.g)

  def apply(self, inputs, axis: int = 0, lse_mode: bool = True) -> torch.Tensor:
    """sinkhorn.py:456-470: the transport applied to ``inputs`` (axis 1: its transpose)."""
    if not lse_mode:
      raise NotImplementedError("apply with lse_mode=False is not supported by ott_amd")
    return self.geom.apply_transport_from_potentials(self.f, self.g, inputs, axis=axis)

  def to_dual_potentials(self, epsilon: Optional[float] = None):
    """sinkhorn.py:480-499: the dual potential functions, soft c-transforms of
    g against (y, b) and of f against (x, a) (linear_problem.py:111-171)."""
    from ott_amd.problems.linear import potentials
    return potentials.from_solution(self.geom, self.f, self.g, self.a, self.b, epsilon=epsilon)

  def transport_cost_at_geom(self, other_geom: PointCloud) -> float:
    """sinkhorn.py:386-411: <P, C'> for another PointCloud's cost C'.

    The reference sums P against the cost's low-rank factors
    (pointcloud.py:298-319); here one fused pass over the pairs forms P_ij and
    C'_ij from the two pairs of clouds and sums P_ij C'_ij in float64, never
    materialising P or C' (``ottx_transport_cost``).
    """
    if not isinstance(other_geom, PointCloud):
      raise NotImplementedError("transport_cost_at_geom needs a PointCloud")
    if tuple(other_geom.shape) != tuple(self.geom.shape):
      raise ValueError(f"other_geom must be {tuple(self.geom.shape)}, got {tuple(other_geom.shape)}")
    return self.geom.transport_cost(self.f, self.g, None if other_geom is self.geom else other_geom)

  @property
  def primal_cost(self) -> float:
    """sinkhorn.py:345-348."""
    return self.transport_cost_at_geom(other_geom=self.geom)

  def reg_ot_cost_grad(self) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Gradients of ``reg_ot_cost`` with respect to x, y, a and b with the
    potentials held fixed: what jax.grad returns through the reference's
    reg_ot_cost, whose f and g are stop_gradient'ed under the default
    ``use_danskin`` (sinkhorn.py:329-335, compute_kl_reg_cost :224-277).

    - x: 2 s sum_j P_ij (x_i - y_j), y: 2 s sum_i P_ij (y_j - x_i), on the GPU
      (``ottx_cost_gradient``, d <= 64);
    - a: f - eps log a - eps + eps sum(b) where a > 0 (0 at zero weights), b
      likewise.

    Balanced problems with an absolute epsilon and a float ``scale_cost``: a
    relative epsilon or a data-dependent scale would add terms through eps or
    s, which this does not include.
    """
    geom = self.geom
    if self.ot_prob.tau_a != 1.0 or self.ot_prob.tau_b != 1.0:
      raise NotImplementedError("reg_ot_cost_grad supports balanced problems only")
    if geom._relative_epsilon is not None or not isinstance(geom._scale_cost, (int, float)):
      raise NotImplementedError("reg_ot_cost_grad needs an absolute epsilon and a float scale_cost")
    if not geom._schedule_is_target_only():
      raise NotImplementedError("reg_ot_cost_grad needs a fixed epsilon")
    return (geom.cost_gradient(self.f, self.g, side=0), geom.cost_gradient(self.f, self.g, side=1),
            *self._weight_grads())

  def _weight_grads(self) -> Tuple[torch.Tensor, torch.Tensor]:
    eps = self.geom.epsilon
    a, b = self.a.double(), self.b.double()
    f, g = self.f.double(), self.g.double()
    zero_a = torch.zeros_like(a)
    zero_b = torch.zeros_like(b)
    ga = torch.where(a > 0, f - eps * torch.log(torch.where(a > 0, a, 1.0)) - eps, zero_a) + eps * b.sum()
    gb = torch.where(b > 0, g - eps * torch.log(torch.where(b > 0, b, 1.0)) - eps, zero_b) + eps * a.sum()
    return ga, gb

  def _potential_marginal_terms(self) -> float:
    """<f, P 1> + <g, 1^T P> with non-finite potentials read as 0."""
    f = self.f.double()
    g = self.g.double()
    safe_f = torch.where(torch.isfinite(f), f, torch.zeros_like(f))
    safe_g = torch.where(torch.isfinite(g), g, torch.zeros_like(g))
    return float((safe_f * self.marginal(1).double()).sum() + (safe_g * self.marginal(0).double()).sum())

  def cost_at_geom(self, other_geom: PointCloud) -> float:
    """sinkhorn.py:474-479: <P, C'> - eps sum(entr(P)).

    With log P_ij = (f_i + g_j - C_ij) / eps, -eps sum(entr(P)) =
    <f, P 1> + <g, 1^T P> - <P, C>, so no matrix is formed.
    """
    return self.transport_cost_at_geom(other_geom) + self._potential_marginal_terms() - self.primal_cost

  @property
  def entropy(self) -> float:
    """sinkhorn.py:512-522."""
    return (self.primal_cost - self._potential_marginal_terms()) / self.geom.epsilon

  @property
  def normalized_entropy(self) -> float:
    """sinkhorn.py:524-529."""
    assert self.ot_prob.is_assignment, "Normalized entropy only valid for assignment problem."
    return self.entropy / math.log(self.geom.shape[0]) - 1.0

  @property
  def diag(self) -> torch.Tensor:
    """sinkhorn.py:531-538: P_ii = exp((f_i + g_i - C_ii) / eps), C_ii the cost
    of x_i against y_i (pointcloud diag_cost)."""
    n, m = self.geom.shape
    assert n == m, "Problem must be square for transport matrix to have a diag."
    diag_cost = self.geom.inv_scale_cost * self.geom.cost_fn(self.geom.x.double(), self.geom.y.double())
    return torch.exp((self.f.double() + self.g.double() - diag_cost) / self.geom.epsilon).float()


class Sinkhorn:
  """Sinkhorn solver with OTT's constructor (sinkhorn.py:718-780).

  Supported on the GPU path: ``lse_mode=True``; any ``threshold``,
  ``norm_error`` (p >= 1), ``inner_iterations``, ``min_iterations``,
  ``max_iterations``; constant ``Momentum`` and adaptive ``Momentum(start)``
  (norm_error 1, start >= 2 inner_iterations); epsilon schedules
  ``Epsilon(target, init, decay)`` with sequential updates; sequential or
  parallel updates (one pass per iteration on the symmetric problems of a
  Sinkhorn divergence, both updates from the previous state otherwise);
  balanced or unbalanced problems; warm starts through ``init`` or an
  ``initializer``;
  ``progress_fn``, called per iteration once each block has run (the solve
  then waits for every block).  Options outside that set raise ``NotImplementedError``;
  there is no silent fallback.  ``implicit_diff`` and ``use_danskin`` only
  affect reverse mode, which this package does not provide.

  Extra (not in OTT): ``use_graph`` replays each block of ``inner_iterations``
  as a hipGraph.
  """

  def __init__(
      self,
      lse_mode: bool = True,
      threshold: float = 1e-3,
      norm_error: int = 1,
      inner_iterations: int = 10,
      min_iterations: int = 0,
      max_iterations: int = 2000,
      momentum: Optional[acceleration.Momentum] = None,
      anderson: Optional[acceleration.AndersonAcceleration] = None,
      parallel_dual_updates: bool = False,
      recenter_potentials: bool = False,
      use_danskin: Optional[bool] = None,
      implicit_diff: Any = True,
      initializer: Optional[init_lib.SinkhornInitializer] = None,
      progress_fn: Optional[Callable] = None,
      use_graph: bool = True,
  ):
    self.lse_mode = lse_mode
    self.threshold = threshold
    self.inner_iterations = inner_iterations
    self.min_iterations = min_iterations
    self.max_iterations = max_iterations
    self._norm_error = norm_error
    self.anderson = anderson
    self.implicit_diff = implicit_diff
    if momentum is not None:
      self.momentum = acceleration.Momentum(momentum.start, momentum.error_threshold,
                                            momentum.value, self.inner_iterations)
    else:
      self.momentum = acceleration.Momentum()
    self.parallel_dual_updates = parallel_dual_updates
    self.recenter_potentials = recenter_potentials
    self.initializer = init_lib.DefaultInitializer() if initializer is None else initializer
    self.progress_fn = progress_fn
    self.use_danskin = (self.implicit_diff is not None) if use_danskin is None else use_danskin
    self.use_graph = use_graph

  @property
  def norm_error(self) -> Tuple[int, ...]:
    return (self._norm_error,)

  @property
  def outer_iterations(self) -> int:
    """sinkhorn.py:964-972 (errors rows)."""
    return int(math.ceil(self.max_iterations / self.inner_iterations))

  def _check_supported(self, ot_prob: linear_problem.LinearProblem) -> None:
    self._check_options()
    if not isinstance(ot_prob.geom, PointCloud):
      raise NotImplementedError("ott_amd accelerates PointCloud geometries only")
    if not ot_prob.geom._schedule_is_target_only() and (
        self.parallel_dual_updates or
        (self.recenter_potentials and ot_prob.tau_a < 1.0 and ot_prob.tau_b < 1.0)):
      # the schedule runs with sequential updates, and not with the recentred
      # unbalanced iteration
      raise NotImplementedError("epsilon schedules (init > 1) are not supported with parallel_dual_updates "
                                "or with recenter_potentials on problems with both taus < 1")
    if self.recenter_potentials and self.max_iterations == 0:
      raise NotImplementedError("recenter_potentials with max_iterations = 0 is not supported")
    if self.parallel_dual_updates and self.recenter_potentials:
      raise NotImplementedError("parallel_dual_updates with recenter_potentials is not supported")

  def _check_options(self) -> None:
    """Solver options the HIP path runs; the others raise NotImplementedError."""
    if not self.lse_mode:
      raise NotImplementedError("ott_amd runs the log-domain (lse_mode=True) solver only")
    if self.anderson is not None:
      raise NotImplementedError("Anderson acceleration is not supported by ott_amd")
    if not self.momentum.is_constant:
      # Momentum(start > 0) (acceleration.py:138-160) reads the L1 errors of the
      # two blocks before ``start``
      # with norm_error != 1 the reference records the first norm's error in
      # the L1 column it adds (sinkhorn.py:1032-1038, 905-931), which is what
      # the weight then reads: the errors of the requested norm
      if self.parallel_dual_updates:
        raise NotImplementedError("adaptive momentum is supported with sequential updates")
      if self.momentum.start < 2 * self.inner_iterations:
        raise NotImplementedError("adaptive momentum needs start >= 2 * inner_iterations")
    if int(self._norm_error) < 1:
      raise ValueError("norm_error must be >= 1")

  def __call__(self, ot_prob: linear_problem.LinearProblem,
               init: Optional[Tuple[Any, Any]] = None, **kwargs: Any) -> SinkhornOutput:
    """sinkhorn.py:782-801."""
    self._check_supported(ot_prob)
    geom = ot_prob.geom
    n, m = geom.shape
    if init is None and getattr(self.initializer, "predicts_g", False):
      # A warm start that predicts g only: the solver derives f0 from it with
      # its own first f-update instead of a separate pass over the geometry.
      g = as_device_f32(self.initializer.predict_g(ot_prob), geom.device)
      f = torch.empty((n,), dtype=torch.float32, device=geom.device)
      return run(ot_prob, self, (f, g), init_mode=_lib.INIT_F_FROM_G)
    if init is None and getattr(self.initializer, "predicts_f", False):
      # A warm start that predicts f only (OTT's learned initializer predicts
      # f too, meta_initializer.py:192-207): its prediction is a fresh tensor
      # that becomes the solve's f; g starts as the DefaultInitializer's, which
      # the first g-update discards but for its -inf entries (geometry.py:297-299).
      # The network is launched first and writes its prediction into the
      # solve's f; the host sets the solve up while the GPU runs the network.
      f = torch.empty((n,), dtype=torch.float32, device=geom.device)
      g = torch.empty((m,), dtype=torch.float32, device=geom.device)
      self.initializer.predict_f(ot_prob, out=f)
      return _prepare(ot_prob, self, (f, g), init_mode=_lib.INIT_F_GIVEN)()
    if init is None and type(self.initializer) is init_lib.DefaultInitializer:
      # zeros, -inf at zero weights (initializers.py:70-99), set by the solve itself
      f = torch.empty((n,), dtype=torch.float32, device=geom.device)
      g = torch.empty((m,), dtype=torch.float32, device=geom.device)
      return run(ot_prob, self, (f, g), init_mode=_lib.INIT_ZEROS)
    if init is None:
      init = self.initializer(ot_prob, lse_mode=self.lse_mode, **kwargs)
    f0, g0 = init
    f = as_device_f32(f0, geom.device).clone()
    g = as_device_f32(g0, geom.device).clone()
    return run(ot_prob, self, (f, g))


def run(ot_prob: linear_problem.LinearProblem, solver: Sinkhorn,
        init: Tuple[torch.Tensor, torch.Tensor], init_mode: int = _lib.INIT_GIVEN) -> SinkhornOutput:
  """sinkhorn.py:1052-1099: the loop, then ``set_cost`` and ``ot_prob``."""
  return _prepare(ot_prob, solver, init, init_mode)()


def _prepare(ot_prob: linear_problem.LinearProblem, solver: Sinkhorn,
             init: Tuple[torch.Tensor, torch.Tensor], init_mode: int):
  """Everything of :func:`run` up to the solve call; returns the call."""
  lib = _lib.load_library()
  geom: PointCloud = ot_prob.geom
  f, g = init
  n, m = geom.shape
  if f.shape != (n,) or g.shape != (m,):
    raise ValueError(f"init potentials must be [{n}] and [{m}]")
  a, b = ot_prob._a, ot_prob._b  # None: uniform, set inside the library
  eps_abs, eps_rel, eps_stat = geom._epsilon_spec()
  sched_init, sched_decay = geom._schedule()
  pb, keep = geom._problem(a=a, b=b, tau_a=ot_prob.tau_a, tau_b=ot_prob.tau_b,
                           epsilon=eps_abs, relative_epsilon=eps_rel, epsilon_statistic=eps_stat)
  opts = _lib.SinkhornOpts(
      threshold=float(solver.threshold), norm_error=int(solver._norm_error),
      inner_iterations=int(solver.inner_iterations),
      min_iterations=int(solver.min_iterations),
      max_iterations=int(solver.max_iterations),
      momentum=float(solver.momentum.value),
      tile_config=tune.lookup(geom.precision, geom.x.shape[1], n, m) if geom._closed_form else -1,
      use_graph=1 if solver.use_graph else 0,
      init_mode=int(init_mode),
      recenter_potentials=1 if solver.recenter_potentials else 0,
      parallel_dual_updates=1 if solver.parallel_dual_updates else 0,
      epsilon_init=sched_init, epsilon_decay=sched_decay,
      momentum_start=int(solver.momentum.start),
      momentum_error_threshold=float(solver.momentum.error_threshold))
  progress = _progress_callback(solver)
  if progress is not None:
    opts.progress_fn = progress.cfunc
  errors = torch.empty((solver.outer_iterations,), dtype=torch.float32, device=geom.device)
  out = _lib.SinkhornOut(f=f.data_ptr(), g=g.data_ptr(), errors=errors.data_ptr())
  nbytes = lib.ottx_sinkhorn_workspace_bytes(ctypes.byref(pb), ctypes.byref(opts))
  if nbytes == 0:
    raise NotImplementedError(f"problem not supported by the HIP kernels: {geom!r}")
  ws = _lib.workspace(nbytes, geom.device)
  stream = _lib.stream_ptr()

  def launch() -> SinkhornOutput:
    _lib.check("ottx_sinkhorn_solve", lib.ottx_sinkhorn_solve(
        ctypes.byref(pb), ctypes.byref(opts), ctypes.byref(out), ws.data_ptr(),
        ws.numel(), stream))
    if progress is not None and progress.error is not None:
      raise progress.error
    geom._set_resolved_epsilon(out.epsilon)
    return SinkhornOutput(
        potentials=(f, g), errors=errors, reg_ot_cost=float(out.reg_ot_cost),
        ot_prob=ot_prob, threshold=float(solver.threshold),
        converged=bool(out.converged), inner_iterations=int(solver.inner_iterations),
        solve_n_iters=int(out.n_iters))

  launch.keep = (keep, progress)  # weights referenced by pb, the callback, alive until the call
  return launch


class ProgressState(NamedTuple):
  """What a ``progress_fn`` receives as the solver state: ``errors`` as the
  reference's ``SinkhornState.errors`` reads at that iteration ([outer, 1],
  host float32; -1 for blocks not run, +inf for iterations without an
  error).  The potentials stay on the GPU and are not part of it."""

  errors: Any


class _Progress:
  """Per-iteration ``progress_fn`` calls (sinkhorn.py:933-937) from the
  solve's per-block callback: block k stands for iterations k inner .. k inner
  + inner - 1, whose state shows +inf in row k until its last iteration, which
  records the block's error (+inf before ``min_iterations``,
  sinkhorn.py:914-931)."""

  def __init__(self, solver: "Sinkhorn"):
    self.fn = solver.progress_fn
    self.inner = int(solver.inner_iterations)
    self.min_it = int(solver.min_iterations)
    self.max_it = int(solver.max_iterations)
    self.errors = np.full((solver.outer_iterations, 1), -1.0, dtype=np.float32)
    self.error: Optional[BaseException] = None
    self.cfunc = _lib.PROGRESS_FN(self._on_block)

  def _on_block(self, user, block: int, err: float) -> None:
    if self.error is not None:
      return
    try:
      for j in range(self.inner):
        it = block * self.inner + j
        last = j == self.inner - 1
        record = last and (it >= self.min_it or it == self.max_it - 1)
        self.errors[block, 0] = np.float32(err) if record else np.inf
        self.fn((it, self.inner, self.max_it, ProgressState(errors=self.errors.copy())))
    except BaseException as e:  # re-raised by the solve's caller
      self.error = e


def _progress_callback(solver: "Sinkhorn") -> Optional[_Progress]:
  return None if solver.progress_fn is None else _Progress(solver)
