"""PointCloud geometry backed by the ottx HIP kernels.

Mirrors ``ott.geometry.pointcloud.PointCloud`` (geometry/pointcloud.py:27-454)
and the parts of ``ott.geometry.geometry.Geometry`` the Sinkhorn solver calls
(the operator protocol, geometry/geometry.py:256-420): ``apply_lse_kernel``,
``update_potential``, ``marginal_from_potentials``, ``epsilon``,
``epsilon_scheduler``, ``shape``, ``dtype``.  The cost matrix is never
materialised: every operator is one fused half-step launch.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any, Literal, Optional, Tuple, Union

import numpy as np
import torch

from ott_amd import _lib
from ott_amd.geometry import costs
from ott_amd.geometry import epsilon_scheduler as eps_scheduler

__all__ = ["PointCloud"]

Precision = Literal["fp32", "bf16_split"]


def as_device_f32(arr, device: torch.device) -> torch.Tensor:
  """Contiguous float32 tensor on ``device`` (numpy and CPU tensors are copied)."""
  if isinstance(arr, np.ndarray):
    arr = torch.from_numpy(np.ascontiguousarray(arr))
  if not isinstance(arr, torch.Tensor):
    arr = torch.as_tensor(arr)
  return arr.to(device=device, dtype=torch.float32).contiguous()


class PointCloud:
  """Two point clouds with the squared-Euclidean cost.

  Args mirror geometry/pointcloud.py:54-72 and geometry/geometry.py:76-89:
    x: ``[n, d]`` points.  y: ``[m, d]`` points, ``None`` means ``x``.
    cost_fn: :class:`~ott_amd.geometry.costs.SqEuclidean` (the default, MFMA
      cross term), or ``PNormP``, ``SqPNorm``, ``EuclideanP``, ``Euclidean``
      (summed from the differences on the vector ALUs, online).
    batch_size: accepted for signature compatibility; the kernels are always
      online (the cost is recomputed per tile and never stored).
    scale_cost: a positive float ``s`` so that ``C = ||x - y||^2 / s``.
    epsilon: float, :class:`Epsilon`, or the multiplier when
      ``relative_epsilon`` is set.
    relative_epsilon: ``"mean"`` or ``"std"`` (both in closed form on the
      GPU, the cost matrix is never formed) or ``None``; with ``epsilon`` also
      ``None`` the epsilon is 0.05 * std of the cost, the reference's default.
    precision: ``"fp32"`` (f32 MFMA cross term) or ``"bf16_split"`` (hi/lo
      bfloat16 operands, f32 accumulation).  Not part of OTT's signature.
  """

  def __init__(
      self,
      x,
      y=None,
      cost_fn: Optional[costs.CostFn] = None,
      batch_size: Optional[int] = None,
      scale_cost: Union[float, str] = 1.0,
      epsilon: Union[float, eps_scheduler.Epsilon, None] = None,
      relative_epsilon: Optional[Literal["mean", "std"]] = None,
      precision: Precision = "fp32",
      **kwargs: Any,
  ):
    if kwargs:
      raise TypeError(f"unexpected PointCloud arguments: {sorted(kwargs)}")
    self.device = _lib.require_device()
    self.x = as_device_f32(x, self.device)
    self.y = self.x if y is None else as_device_f32(y, self.device)
    if self.x.ndim != 2 or self.y.ndim != 2 or self.x.shape[1] != self.y.shape[1]:
      raise ValueError(f"x and y must be [n, d] and [m, d], got {tuple(self.x.shape)} "
                       f"and {tuple(self.y.shape)}")
    self.cost_fn = costs.SqEuclidean() if cost_fn is None else cost_fn
    # the kernels evaluate factor * cost(kind, p) (factor 0.5 for PNormP(2)
    # and SqPNorm(2), which run as half the squared Euclidean cost)
    self._cost_kind, self._cost_p, self._cost_factor = self.cost_fn.kernel_spec()
    if batch_size is not None:
      assert batch_size > 0, f"`batch_size={batch_size}` must be positive."
    self._batch_size = batch_size
    self._scale_cost = scale_cost
    self._epsilon_init = epsilon
    self._relative_epsilon = relative_epsilon
    if precision not in _lib.PRECISIONS:
      raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
    self.precision = precision
    self._eps_cache: Optional[eps_scheduler.Epsilon] = None
    self._mean_cache: Optional[float] = None
    self._std_cache: Optional[float] = None
    self._inv_scale_cache: Optional[float] = None
    self._median_cache: Optional[float] = None
    self._summary_cache: Optional[Tuple[float, float, float]] = None

  # -- shape / dtype ---------------------------------------------------------
  @property
  def shape(self) -> Tuple[int, int]:
    return self.x.shape[0], self.y.shape[0]

  @property
  def dtype(self) -> torch.dtype:
    return self.x.dtype

  @property
  def is_squared_euclidean(self) -> bool:
    return isinstance(self.cost_fn, costs.SqEuclidean)

  @property
  def _closed_form(self) -> bool:
    """The cost is (a multiple of) the squared Euclidean one: MFMA half-steps,
    closed-form mean and std of the cost matrix."""
    return self._cost_kind == costs.COST_SQEUCLIDEAN

  @property
  def is_online(self) -> bool:
    return self._batch_size is not None

  @property
  def batch_size(self) -> Optional[int]:
    if self._batch_size is None:
      return None
    n, m = self.shape
    return min(n, m, self._batch_size)

  @property
  def cost_rank(self) -> int:
    return self.x.shape[1]

  # -- cost scaling and epsilon ---------------------------------------------
  @property
  def inv_scale_cost(self) -> float:
    """pointcloud.py:342-383.

    Floats scale directly.  "mean" uses the closed-form mean of the unscaled
    cost, "max_norm" / "max_bound" the largest squared norms, "max_cost" the
    largest pair cost and "median" the median entry (jnp.median of the
    unscaled matrix, as the reference computes for clouds that are not online;
    an exact selection over the pairs on the GPU, ``ottx_cost_median``), each
    computed once and kept.
    """
    s = self._scale_cost
    if isinstance(s, (int, float, np.floating)) and not isinstance(s, bool):
      if not float(s) > 0.0:
        raise ValueError(f"scale_cost must be positive, got {s}")
      return 1.0 / float(s)
    if isinstance(s, str) and s in ("mean", "max_cost", "max_norm", "max_bound", "median"):
      if self._inv_scale_cache is None:
        v = self._cost_summary(s)
        self._inv_scale_cache = math.inf if v == 0.0 else 1.0 / v
      return self._inv_scale_cache
    raise ValueError(f"Scaling {s} not implemented.")

  @property
  def _kernel_scale(self) -> float:
    """The s the kernels multiply their cost by: inv_scale_cost times the
    factor of a cost that runs as a multiple of the squared Euclidean one."""
    return self.inv_scale_cost * self._cost_factor

  def _cost_summary(self, mode: str) -> float:
    """The unscaled quantity a scale_cost mode divides by (pointcloud.py:343-377)."""
    lib = _lib.load_library()
    n, m = self.shape
    d = self.x.shape[1]
    if mode == "median":
      return self._unscaled_median()
    if not self._closed_form and mode in ("mean", "max_cost"):
      total, _, c_max = self._cost_summary_values()
      return c_max if mode == "max_cost" else total / (float(n) * float(m))
    if mode == "mean":
      nbytes = lib.ottx_mean_cost_workspace_bytes(n, m, d)
      ws = _lib.workspace(nbytes, self.device)
      out = ctypes.c_double(0.0)
      _lib.check("ottx_mean_cost", lib.ottx_mean_cost(
          n, m, d, self.x.data_ptr(), self.y.data_ptr(), self._cost_factor, ctypes.byref(out), ws.data_ptr(),
          ws.numel(), _lib.stream_ptr()))
      return float(out.value)
    ws = _lib.workspace(256, self.device)
    out = (ctypes.c_double * 3)()
    _lib.check("ottx_cost_maxima", lib.ottx_cost_maxima(
        n, m, d, self.x.data_ptr(), self.y.data_ptr(), 1 if mode == "max_cost" else 0, out, ws.data_ptr(),
        ws.numel(), _lib.stream_ptr()))
    if not self.is_squared_euclidean and mode in ("max_norm", "max_bound"):
      raise NotImplementedError(f"scale_cost={mode!r} needs the squared-Euclidean cost (cost_fn.norm)")
    x_max, y_max, c_max = float(out[0]), float(out[1]), float(out[2])
    if mode == "max_cost":
      return self._cost_factor * c_max
    if mode == "max_norm":
      return max(x_max, y_max)
    # max_bound for the squared-Euclidean cost
    return x_max + y_max + 2.0 * math.sqrt(x_max * y_max)

  def _cost_problem(self) -> _lib.Problem:
    """The clouds and the cost for the summary entry points (no epsilon)."""
    n, m = self.shape
    return _lib.Problem(n=n, m=m, d=self.x.shape[1], precision=_lib.PRECISIONS[self.precision],
                        x=self.x.data_ptr(), y=self.y.data_ptr(), a=None, b=None, epsilon=1.0,
                        inv_scale_cost=1.0, tau_a=1.0, tau_b=1.0, relative_epsilon=0.0,
                        epsilon_statistic=_lib.EPS_STAT_MEAN, cost=self._cost_kind, cost_p=self._cost_p)

  def _unscaled_median(self) -> float:
    if self._median_cache is None:
      lib = _lib.load_library()
      ws = _lib.workspace(lib.ottx_cost_median_workspace_bytes(), self.device)
      out = ctypes.c_double(0.0)
      pb = self._cost_problem()
      _lib.check("ottx_cost_median", lib.ottx_cost_median(
          ctypes.byref(pb), ctypes.byref(out), ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
      self._median_cache = self._cost_factor * float(out.value)
    return self._median_cache

  def _cost_summary_values(self) -> Tuple[float, float, float]:
    """(sum C, sum C^2, max C) of the unscaled cost matrix, one GPU pass."""
    if self._summary_cache is None:
      lib = _lib.load_library()
      n, m = self.shape
      ws = _lib.workspace(lib.ottx_cost_summary_workspace_bytes(n, m), self.device)
      out = (ctypes.c_double * 3)()
      pb = self._cost_problem()
      _lib.check("ottx_cost_summary", lib.ottx_cost_summary(
          ctypes.byref(pb), out, ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
      f = self._cost_factor
      self._summary_cache = (f * float(out[0]), f * f * float(out[1]), f * float(out[2]))
    return self._summary_cache

  @property
  def median_cost_matrix(self) -> float:
    """geometry.py:106-109: the median of the (scaled) cost matrix, exact, on
    the GPU without forming the matrix."""
    return self.inv_scale_cost * self._unscaled_median()

  @property
  def mean_cost_matrix(self) -> float:
    """geometry.py:112-117 via pointcloud.py:170-186, closed form on the GPU
    for the squared Euclidean cost, one pass over the pairs for the others."""
    if self._mean_cache is None and not self._closed_form:
      n, m = self.shape
      self._mean_cache = self.inv_scale_cost * self._cost_summary_values()[0] / (float(n) * float(m))
    if self._mean_cache is None:
      lib = _lib.load_library()
      n, m = self.shape
      d = self.x.shape[1]
      nbytes = lib.ottx_mean_cost_workspace_bytes(n, m, d)
      ws = _lib.workspace(nbytes, self.device)
      out = ctypes.c_double(0.0)
      _lib.check("ottx_mean_cost", lib.ottx_mean_cost(
          n, m, d, self.x.data_ptr(), self.y.data_ptr(), self._kernel_scale,
          ctypes.byref(out), ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
      self._mean_cache = float(out.value)
    return self._mean_cache

  @property
  def std_cost_matrix(self) -> float:
    """geometry.py:119-136: sqrt(relu(mean(C^2) - mean(C)^2)), in closed form
    on the GPU from the clouds' moments for the squared Euclidean cost (the
    reference forms C^2), from float64 sums over the pairs for the others."""
    if self._std_cache is None and not self._closed_form:
      n, m = self.shape
      s1, s2, _ = self._cost_summary_values()
      nm = float(n) * float(m)
      var = s2 / nm - (s1 / nm) ** 2
      self._std_cache = self.inv_scale_cost * math.sqrt(max(var, 0.0))
    if self._std_cache is None:
      lib = _lib.load_library()
      n, m = self.shape
      d = self.x.shape[1]
      nbytes = lib.ottx_cost_moments_workspace_bytes(n, m, d)
      if nbytes == 0:
        raise NotImplementedError(f"std_cost_matrix supports d <= 256, got d={d}")
      ws = _lib.workspace(nbytes, self.device)
      out = (ctypes.c_double * 2)()
      _lib.check("ottx_cost_moments", lib.ottx_cost_moments(
          n, m, d, self.x.data_ptr(), self.y.data_ptr(), self._kernel_scale, out, ws.data_ptr(), ws.numel(),
          _lib.stream_ptr()))
      if self._mean_cache is None:
        self._mean_cache = float(out[0])
      self._std_cache = float(out[1])
    return self._std_cache

  def _relative_statistic(self) -> Optional[int]:
    """The cost statistic the epsilon multiplies (None: an absolute epsilon)
    and, through it, geometry.py:147-171's branches."""
    init = self._epsilon_init
    if isinstance(init, eps_scheduler.Epsilon):
      return None
    if self._relative_epsilon is None:
      return None if init is not None else _lib.EPS_STAT_STD
    if self._relative_epsilon == "std":
      return _lib.EPS_STAT_STD
    if self._relative_epsilon == "mean":
      return _lib.EPS_STAT_MEAN
    raise ValueError(f"Invalid relative epsilon: {self._relative_epsilon}.")

  @property
  def epsilon_scheduler(self) -> eps_scheduler.Epsilon:
    """geometry.py:147-171."""
    if self._eps_cache is not None:
      return self._eps_cache
    init = self._epsilon_init
    stat = self._relative_statistic()
    if isinstance(init, eps_scheduler.Epsilon):
      sched = init
    elif stat is None:
      sched = eps_scheduler.Epsilon(float(init))
    else:
      mult = eps_scheduler.DEFAULT_EPSILON_SCALE if init is None else float(init)
      scale = self.mean_cost_matrix if stat == _lib.EPS_STAT_MEAN else self.std_cost_matrix
      sched = eps_scheduler.Epsilon(mult * scale)
    self._eps_cache = sched
    return sched

  @property
  def epsilon(self) -> float:
    return float(self.epsilon_scheduler.target)

  def _epsilon_spec(self) -> Tuple[float, float, int]:
    """(absolute epsilon, relative multiplier, statistic) for the C ABI.

    A not-yet-resolved relative epsilon (``"mean"``, ``"std"`` or the default
    0.05 * std) is passed as a multiplier so the solve resolves it on the
    device without a host round trip.
    """
    stat = self._relative_statistic()
    if self._eps_cache is None and stat is not None and self._closed_form:
      mult = (eps_scheduler.DEFAULT_EPSILON_SCALE if self._epsilon_init is None
              else float(self._epsilon_init))
      return 0.0, mult, stat
    return self.epsilon, 0.0, _lib.EPS_STAT_MEAN

  def _set_resolved_epsilon(self, eps: float) -> None:
    if self._eps_cache is None:
      self._eps_cache = eps_scheduler.Epsilon(float(eps))

  def _schedule_is_target_only(self) -> bool:
    init = self._epsilon_init
    return not isinstance(init, eps_scheduler.Epsilon) or init.is_target_only

  def _schedule(self) -> Tuple[float, float]:
    """(init, decay) of the epsilon schedule, (1, 1) for a fixed epsilon."""
    init = self._epsilon_init
    if isinstance(init, eps_scheduler.Epsilon) and not init.is_target_only:
      return float(init.init), float(init.decay)
    return 1.0, 1.0

  # -- divergences ------------------------------------------------------------
  @classmethod
  def prepare_divergences(cls, x, y, static_b: bool = False, **kwargs: Any) -> Tuple["PointCloud", ...]:
    """pointcloud.py:226-237: the geometries (x, y), (x, x) and, unless
    ``static_b``, (y, y).  The symmetric ones hold one cloud on both sides (the
    same device buffer), which the solver's parallel updates require."""
    geoms = [cls(x, y, **kwargs), cls(x, None, **kwargs)]
    if not static_b:
      geoms.append(cls(y, None, **kwargs))
    return tuple(geoms)

  def copy_epsilon(self, other: "PointCloud") -> "PointCloud":
    """geometry.py:248-254: this geometry with ``other``'s epsilon (resolved)."""
    return type(self)(self.x, None if self.y is self.x else self.y, cost_fn=self.cost_fn,
                      batch_size=self._batch_size, scale_cost=self._scale_cost,
                      epsilon=other.epsilon_scheduler, relative_epsilon=other._relative_epsilon,
                      precision=self.precision)

  # -- the operator protocol -------------------------------------------------
  def _problem(self, a: Optional[torch.Tensor] = None, b: Optional[torch.Tensor] = None,
               tau_a: float = 1.0, tau_b: float = 1.0, epsilon: Optional[float] = None,
               relative_epsilon: float = 0.0,
               epsilon_statistic: int = _lib.EPS_STAT_MEAN) -> Tuple[_lib.Problem, Tuple]:
    n, m = self.shape
    # None weights are passed as NULL: the library uses uniform 1/n, 1/m
    pb = _lib.Problem(
        n=n, m=m, d=self.x.shape[1], precision=_lib.PRECISIONS[self.precision],
        x=self.x.data_ptr(), y=self.y.data_ptr(), a=_lib.ptr(a), b=_lib.ptr(b),
        epsilon=float(self.epsilon if epsilon is None else epsilon),
        inv_scale_cost=self._kernel_scale, tau_a=float(tau_a), tau_b=float(tau_b),
        relative_epsilon=float(relative_epsilon), epsilon_statistic=int(epsilon_statistic),
        cost=self._cost_kind, cost_p=self._cost_p)
    return pb, (a, b)

  def _op(self, op: int, f, g, eps: float, axis: int,
          log_marginal: Optional[torch.Tensor] = None) -> torch.Tensor:
    if axis not in (0, 1):
      raise ValueError(f"axis must be 0 or 1, got {axis}")
    lib = _lib.load_library()
    f = as_device_f32(f, self.device)
    g = as_device_f32(g, self.device)
    n, m = self.shape
    if f.shape != (n,) or g.shape != (m,):
      raise ValueError(f"potentials must be [{n}] and [{m}]")
    pb, keep = self._problem(epsilon=eps)
    out = torch.empty((m if axis == 0 else n,), dtype=torch.float32, device=self.device)
    lm = None if log_marginal is None else as_device_f32(log_marginal, self.device)
    nbytes = lib.ottx_op_workspace_bytes(ctypes.byref(pb))
    ws = _lib.workspace(nbytes, self.device)
    _lib.check("ottx_pointcloud_op", lib.ottx_pointcloud_op(
        ctypes.byref(pb), op, axis, f.data_ptr(), g.data_ptr(), _lib.ptr(lm),
        float(eps), out.data_ptr(), ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
    del keep
    return out

  def apply_lse_kernel(self, f, g, eps: float, vec=None, axis: int = 0):
    """geometry.py:260-299 (vec=None) -> ``(eps*LSE(...) - h, sign)``."""
    if vec is not None:
      raise NotImplementedError("apply_lse_kernel with vec is not supported yet")
    res = self._op(_lib.OP_APPLY_LSE, f, g, float(eps), axis)
    return res, torch.ones((1,), dtype=torch.float32, device=self.device)

  def update_potential(self, f, g, log_marginal, iteration: Optional[int] = None,
                       axis: int = 0) -> torch.Tensor:
    """geometry.py:376-398."""
    eps = self.epsilon_scheduler(iteration)
    return self._op(_lib.OP_UPDATE, f, g, float(eps), axis, log_marginal=log_marginal)

  def marginal_from_potentials(self, f, g, axis: int = 0) -> torch.Tensor:
    """geometry.py:326-349 (target epsilon)."""
    return self._op(_lib.OP_MARGINAL, f, g, self.epsilon, axis)

  # -- transport from potentials (output side) --------------------------------
  def transport_from_potentials(self, f, g) -> torch.Tensor:
    """geometry.py:361-365: the dense [n, m] transport matrix (target epsilon)."""
    lib = _lib.load_library()
    n, m = self.shape
    f = as_device_f32(f, self.device)
    g = as_device_f32(g, self.device)
    out = torch.empty((n, m), dtype=torch.float32, device=self.device)
    pb, keep = self._problem(epsilon=self.epsilon)
    _lib.check("ottx_transport_matrix", lib.ottx_transport_matrix(
        ctypes.byref(pb), f.data_ptr(), g.data_ptr(), float(self.epsilon), out.data_ptr(), _lib.stream_ptr()))
    del keep
    return out

  def transport_cost(self, f, g, other: Optional["PointCloud"] = None) -> float:
    """sum_ij P_ij C'_ij with P = exp((f + g - C) / eps) (target epsilon) and C'
    the cost of ``other`` (this geometry's own cost when None): the quantity
    of sinkhorn.py:345-411, in one pass over the pairs."""
    lib = _lib.load_library()
    n, m = self.shape
    f = as_device_f32(f, self.device)
    g = as_device_f32(g, self.device)
    if f.shape != (n,) or g.shape != (m,):
      raise ValueError(f"potentials must be [{n}] and [{m}]")
    pb, keep = self._problem(epsilon=self.epsilon)
    nbytes = lib.ottx_transport_cost_workspace_bytes(n, m)
    ws = _lib.workspace(nbytes, self.device)
    out = ctypes.c_double(0.0)
    if other is None:
      xo = yo = None
      d_o, s_o = 0, 0.0
    else:
      if not other._closed_form:
        raise NotImplementedError("transport_cost_at_geom: another geometry's cost must be the squared "
                                  "Euclidean one (or a multiple of it)")
      xo, yo = other.x, other.y
      d_o, s_o = int(xo.shape[1]), float(other._kernel_scale)
    _lib.check("ottx_transport_cost", lib.ottx_transport_cost(
        ctypes.byref(pb), f.data_ptr(), g.data_ptr(), float(self.epsilon), _lib.ptr(xo), _lib.ptr(yo), d_o, s_o,
        ctypes.byref(out), ws.data_ptr(), nbytes, _lib.stream_ptr()))
    del keep, ws
    return float(out.value)

  def cost_gradient(self, f, g, side: int = 0) -> torch.Tensor:
    """sum_j P_ij dC_ij/dx_i (side 0, [n, d]) or sum_i P_ij dC_ij/dy_j (side 1,
    [m, d]) with P = exp((f + g - C) / eps) at the target epsilon: the gradient
    of the coupling's total mass term of reg_ot_cost (sinkhorn.py:224-277)
    with the potentials held fixed."""
    lib = _lib.load_library()
    n, m = self.shape
    f = as_device_f32(f, self.device)
    g = as_device_f32(g, self.device)
    if f.shape != (n,) or g.shape != (m,):
      raise ValueError(f"potentials must be [{n}] and [{m}]")
    pb, keep = self._problem(epsilon=self.epsilon)
    nbytes = lib.ottx_cost_gradient_workspace_bytes(ctypes.byref(pb), int(side))
    if nbytes == 0:
      raise NotImplementedError("cost_gradient supports d <= 64 (side 0 or 1)")
    ws = _lib.workspace(nbytes, self.device)
    out = torch.empty(((n if side == 0 else m), self.x.shape[1]), dtype=torch.float32, device=self.device)
    _lib.check("ottx_cost_gradient", lib.ottx_cost_gradient(
        ctypes.byref(pb), f.data_ptr(), g.data_ptr(), float(self.epsilon), int(side), out.data_ptr(),
        ws.data_ptr(), nbytes, _lib.stream_ptr()))
    del keep, ws  # stream-ordered reuse by the caching allocator
    return out

  def apply_transport_from_potentials(self, f, g, vec, axis: int = 0) -> torch.Tensor:
    """geometry.py:445-486: P^T vec (axis 0, vec over the n rows) or P vec
    (axis 1, vec over the m columns), P = exp((f + g - C) / eps), for one
    vector or a batch [k, n or m], signs included.

    ``ottx_pointcloud_apply``: one log-sum-exp pass over the pairs, then one
    launch per 16 vectors that forms the weights on MFMA and multiplies them
    with the vectors on fp32 MFMA (no host synchronisation).
    """
    if axis not in (0, 1):
      raise ValueError(f"axis must be 0 or 1, got {axis}")
    lib = _lib.load_library()
    f = as_device_f32(f, self.device)
    g = as_device_f32(g, self.device)
    vec = as_device_f32(vec, self.device)
    single = vec.ndim == 1
    vecs = (vec[None] if single else vec).contiguous()
    n, m = self.shape
    if f.shape != (n,) or g.shape != (m,):
      raise ValueError(f"potentials must be [{n}] and [{m}]")
    n_in, n_out = (n, m) if axis == 0 else (m, n)
    if vecs.ndim != 2 or vecs.shape[1] != n_in:
      raise ValueError(f"vec must be [{n_in}] or [k, {n_in}] for axis={axis}")
    k = vecs.shape[0]
    out = torch.empty((k, n_out), dtype=torch.float32, device=self.device)
    if k == 0:
      return out[0] if single else out
    pb, keep = self._problem(epsilon=self.epsilon)
    nbytes = lib.ottx_apply_workspace_bytes(ctypes.byref(pb), int(axis))
    if nbytes == 0:
      raise NotImplementedError("apply_transport_from_potentials: unsupported problem shape")
    ws = _lib.workspace(nbytes, self.device)
    _lib.check("ottx_pointcloud_apply", lib.ottx_pointcloud_apply(
        ctypes.byref(pb), int(axis), f.data_ptr(), g.data_ptr(), vecs.data_ptr(), int(k), float(self.epsilon), 0,
        out.data_ptr(), ws.data_ptr(), nbytes, _lib.stream_ptr()))
    del keep, ws  # stream-ordered reuse by the caching allocator
    return out[0] if single else out

  def transport_operator(self, f, g) -> "TransportOperator":
    """P and P^T as repeatable operators at fixed potentials (iterative
    solvers: the clouds are prepared and each axis' log-sum-exp computed once)."""
    return TransportOperator(self, f, g)

  def low_rank_cost_factors(self) -> Tuple[torch.Tensor, torch.Tensor]:
    """pointcloud.py:298-319 (``to_LRCGeometry`` of the squared-Euclidean cost):
    C = cost_1 cost_2^T with cost_1 = [|x|^2, 1, -2 x] s and cost_2 = [1, |y|^2, y],
    here on points centred by the mean of both clouds (C is unchanged by a
    common shift, and the terms cancel less)."""
    if not self._closed_form:
      raise NotImplementedError("low_rank_cost_factors needs the squared-Euclidean cost (or a multiple of it)")
    mu = 0.5 * (self.x.double().mean(0) + self.y.double().mean(0))
    x = self.x.double() - mu
    y = self.y.double() - mu
    s = self._kernel_scale
    ones_x = torch.ones((x.shape[0], 1), dtype=torch.float64, device=self.device)
    ones_y = torch.ones((y.shape[0], 1), dtype=torch.float64, device=self.device)
    cost_1 = torch.cat([(x * x).sum(1, keepdim=True), ones_x, -2.0 * x], dim=1) * s
    cost_2 = torch.cat([ones_y, (y * y).sum(1, keepdim=True), y], dim=1)
    return cost_1, cost_2

  def __repr__(self) -> str:
    n, m = self.shape
    return (f"PointCloud(n={n}, m={m}, d={self.x.shape[1]}, precision={self.precision!r})")


class TransportOperator:
  """v -> P^T v (axis 0) and v -> P v (axis 1), P = exp((f + g - C) / eps) at
  the geometry's epsilon, for fixed potentials: the workspace keeps the
  prepared clouds and each axis' column log-sum-exp between calls
  (``ottx_pointcloud_apply`` with OTTX_APPLY_PREPARED / OTTX_APPLY_LSE_READY)."""

  def __init__(self, geom: PointCloud, f, g):
    self.geom = geom
    self.f = as_device_f32(f, geom.device).contiguous()
    self.g = as_device_f32(g, geom.device).contiguous()
    n, m = geom.shape
    if self.f.shape != (n,) or self.g.shape != (m,):
      raise ValueError(f"potentials must be [{n}] and [{m}]")
    lib = _lib.load_library()
    self._pb, self._keep = geom._problem(epsilon=geom.epsilon)
    self._nbytes = max(lib.ottx_apply_workspace_bytes(ctypes.byref(self._pb), 0),
                       lib.ottx_apply_workspace_bytes(ctypes.byref(self._pb), 1))
    if self._nbytes == 0:
      raise NotImplementedError("transport operator: unsupported problem shape")
    self._ws = _lib.workspace(self._nbytes, geom.device)
    self._prepared = False
    self._lse_ready = [False, False]

  def __call__(self, vec, axis: int) -> torch.Tensor:
    lib = _lib.load_library()
    vec = as_device_f32(vec, self.geom.device)
    single = vec.ndim == 1
    vecs = (vec[None] if single else vec).contiguous()
    n, m = self.geom.shape
    n_in, n_out = (n, m) if axis == 0 else (m, n)
    if vecs.ndim != 2 or vecs.shape[1] != n_in:
      raise ValueError(f"vec must be [{n_in}] or [k, {n_in}] for axis={axis}")
    out = torch.empty((vecs.shape[0], n_out), dtype=torch.float32, device=self.geom.device)
    flags = (_lib.APPLY_PREPARED if self._prepared else 0) | (_lib.APPLY_LSE_READY if self._lse_ready[axis] else 0)
    _lib.check("ottx_pointcloud_apply", lib.ottx_pointcloud_apply(
        ctypes.byref(self._pb), int(axis), self.f.data_ptr(), self.g.data_ptr(), vecs.data_ptr(),
        int(vecs.shape[0]), float(self.geom.epsilon), flags, out.data_ptr(), self._ws.data_ptr(), self._nbytes,
        _lib.stream_ptr()))
    self._prepared = True
    self._lse_ready[axis] = True
    return out[0] if single else out
