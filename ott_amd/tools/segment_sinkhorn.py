"""Batched Sinkhorn over segments of concatenated point clouds.

Mirrors ``ott.tools.segment_sinkhorn.segment_sinkhorn``
(tools/segment_sinkhorn.py:24-138).  OTT pads every segment to
``max_measure_size`` with zero-weight zero vectors and ``vmap``s one solve per
segment; the batched while-loop keeps each problem's own convergence.  Here the
segments stay ragged: ``ottx_sinkhorn_batched_solve`` runs every problem's
half-steps in shared launches (one work table of (problem, column tile, row
split) per direction) and a problem that has converged drops out of later
launches.  The only effect padding has in OTT besides layout -- a relative
epsilon computed over the padded clouds -- is reproduced through the padded
row counts (``mean_rows_x`` / ``mean_rows_y``), so the costs match OTT's.
"""

from __future__ import annotations

import ctypes
from types import MappingProxyType
from typing import Any, Mapping, NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from ott_amd import _lib
from ott_amd.geometry import costs, epsilon_scheduler as eps_scheduler, segment
from ott_amd.geometry.pointcloud import as_device_f32
from ott_amd.solvers.linear import sinkhorn

__all__ = ["segment_sinkhorn", "segment_sinkhorn_outputs", "SegmentSinkhornOutput"]


class SegmentSinkhornOutput(NamedTuple):
  """Per-segment solver outputs (every SinkhornOutput field, ragged)."""
  f: torch.Tensor            # [sum n_k], rows grouped by segment
  g: torch.Tensor            # [sum m_k]
  errors: torch.Tensor       # [K, outer]
  reg_ot_cost: torch.Tensor  # [K] float64
  converged: torch.Tensor    # [K] bool
  n_iters: torch.Tensor      # [K] int
  epsilon: torch.Tensor      # [K] float64
  x_offsets: np.ndarray      # [K + 1]
  y_offsets: np.ndarray      # [K + 1]
  n_blocks: int
  a: Optional[torch.Tensor] = None  # [sum n_k] the weights the solve used
  b: Optional[torch.Tensor] = None  # [sum m_k]


def _epsilon_args(epsilon, relative_epsilon) -> Tuple[float, float]:
  """(absolute epsilon, relative multiplier) of PointCloud(epsilon, relative_epsilon)."""
  if isinstance(epsilon, eps_scheduler.Epsilon):
    if not epsilon.is_target_only:
      raise NotImplementedError("epsilon schedules with init > 1 are not supported by the batched solve")
    epsilon = epsilon.target
  if relative_epsilon == "mean":
    return 0.0, (eps_scheduler.DEFAULT_EPSILON_SCALE if epsilon is None else float(epsilon))
  if relative_epsilon is None and epsilon is not None:
    return float(epsilon), 0.0
  raise NotImplementedError("segment_sinkhorn supports a float epsilon or relative_epsilon='mean' "
                            f"(got epsilon={epsilon!r}, relative_epsilon={relative_epsilon!r})")


def segment_sinkhorn_outputs(
    x, y, num_segments: Optional[int] = None, max_measure_size: Optional[int] = None,
    cost_fn: Optional[costs.CostFn] = None, segment_ids_x=None, segment_ids_y=None,
    indices_are_sorted: bool = False, num_per_segment_x: Optional[Sequence[int]] = None,
    num_per_segment_y: Optional[Sequence[int]] = None, weights_x=None, weights_y=None,
    sinkhorn_kwargs: Mapping[str, Any] = MappingProxyType({}), tau_a: float = 1.0, tau_b: float = 1.0,
    padded_statistics: bool = True, tile_config: int = -1, epsilons=None, **kwargs: Any) -> SegmentSinkhornOutput:
  """All outputs of the batched solve (see ``segment_sinkhorn``).

  ``padded_statistics=True`` (OTT's behaviour) computes a relative epsilon
  over each segment padded to ``max_measure_size`` rows; ``False`` uses each
  segment's own points, i.e. what one ``PointCloud`` per segment would give.
  ``tile_config`` (beyond OTT) forces one of ``ottx_tile_configs``' variants;
  -1 takes the autotuner's record for the batch's size class
  (``tune.lookup_batch``), else the variant with the least row padding.
  ``epsilons`` (beyond OTT): an absolute epsilon per segment, which replaces
  ``epsilon`` / ``relative_epsilon`` (a Sinkhorn divergence shares the
  (x, y) terms' epsilons with the symmetric ones).
  """
  if cost_fn is not None and not isinstance(cost_fn, costs.SqEuclidean):
    raise NotImplementedError("ott_amd accelerates the squared-Euclidean cost only")
  precision = kwargs.pop("precision", "fp32")
  epsilon = kwargs.pop("epsilon", None)
  relative_epsilon = kwargs.pop("relative_epsilon", None)
  scale_cost = kwargs.pop("scale_cost", 1.0)
  if kwargs.pop("batch_size", None) is not None:
    pass  # online evaluation is how the kernels always work
  if kwargs:
    raise TypeError(f"unexpected PointCloud arguments: {sorted(kwargs)}")
  if isinstance(scale_cost, str):
    raise NotImplementedError(f"scale_cost={scale_cost!r} is not supported by ott_amd yet")
  if not scale_cost > 0:
    raise ValueError(f"scale_cost must be positive, got {scale_cost}")
  if precision not in _lib.PRECISIONS:
    raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}")
  solver = sinkhorn.Sinkhorn(**dict(sinkhorn_kwargs))
  solver._check_options()
  if solver.parallel_dual_updates:
    raise NotImplementedError("parallel_dual_updates is not supported by the batched solve")
  if solver.recenter_potentials:
    raise NotImplementedError("recenter_potentials is not supported by the batched solve")
  if not solver.momentum.is_constant:
    raise NotImplementedError("adaptive momentum is not supported by the batched solve")
  if solver.initializer is not None and not isinstance(solver.initializer, sinkhorn.init_lib.DefaultInitializer):
    raise NotImplementedError("segment_sinkhorn starts every problem from zero potentials")
  eps_abs, eps_rel = _epsilon_args(epsilon, relative_epsilon) if epsilons is None else (0.0, 0.0)

  device = _lib.require_device()
  x = as_device_f32(x, device)
  y = as_device_f32(y, device)
  if x.ndim != 2 or y.ndim != 2 or x.shape[1] != y.shape[1]:
    raise ValueError("x and y must be [num_x, d] and [num_y, d]")
  use_ids = segment_ids_x is not None
  if use_ids and segment_ids_y is None:
    raise ValueError("segment_ids_y is required with segment_ids_x")
  if use_ids and max_measure_size is None:
    raise ValueError("Please specify `max_measure_size`.")
  wx = None if weights_x is None else as_device_f32(weights_x, device)
  wy = None if weights_y is None else as_device_f32(weights_y, device)
  xs, a, x_off = segment.contiguous_segments(x, wx, num_segments, segment_ids_x, indices_are_sorted,
                                             num_per_segment_x, default_weights=False)
  ys, b, y_off = segment.contiguous_segments(y, wy, num_segments, segment_ids_y, indices_are_sorted,
                                             num_per_segment_y, default_weights=False)
  # absent weights: the solve writes 1 / n_k (segment.py:100-104) into these
  flags = 0
  if a is None:
    a = torch.empty(xs.shape[0], dtype=torch.float32, device=device)
    flags |= _lib.BATCH_UNIFORM_A
  if b is None:
    b = torch.empty(ys.shape[0], dtype=torch.float32, device=device)
    flags |= _lib.BATCH_UNIFORM_B
  K = len(x_off) - 1
  if len(y_off) - 1 != K:
    raise ValueError("x and y must hold the same number of segments")
  nx, ny = np.diff(x_off), np.diff(y_off)
  if max_measure_size is None:
    max_measure_size = int(max(nx.max(), ny.max()))
  # segment.py:98: the padded size is capped at the number of rows
  px, py = min(max_measure_size, x.shape[0]), min(max_measure_size, y.shape[0])
  if nx.max() > px or ny.max() > py:
    raise ValueError(f"a segment holds more than max_measure_size={max_measure_size} points "
                     "(OTT would silently drop the excess)")

  x_off_c = (ctypes.c_int64 * (K + 1))(*x_off.tolist())
  y_off_c = (ctypes.c_int64 * (K + 1))(*y_off.tolist())
  eps_c = None
  if epsilons is not None:
    eps_list = [float(e) for e in epsilons]
    if len(eps_list) != K or not all(e > 0.0 for e in eps_list):
      raise ValueError(f"epsilons must hold {K} positive values")
    eps_c = (ctypes.c_double * K)(*eps_list)
  elif eps_rel == 0.0:
    eps_c = (ctypes.c_double * K)(*([eps_abs] * K))
  bt = _lib.Batch(count=K, d=x.shape[1], precision=_lib.PRECISIONS[precision], flags=flags, x=xs.data_ptr(),
                  y=ys.data_ptr(), a=a.data_ptr(), b=b.data_ptr(), x_off=x_off_c, y_off=y_off_c,
                  epsilon=eps_c, relative_epsilon=eps_rel,
                  mean_rows_x=float(px) if eps_rel > 0 and padded_statistics else 0.0,
                  mean_rows_y=float(py) if eps_rel > 0 and padded_statistics else 0.0,
                  inv_scale_cost=1.0 / float(scale_cost), tau_a=float(tau_a), tau_b=float(tau_b))
  if tile_config == -1:
    from ott_amd import tune
    tile_config = tune.lookup_batch(precision, int(x.shape[1]), K, float(np.sum(nx.astype(np.float64) * ny)))
  opts = _lib.SinkhornOpts(
      threshold=float(solver.threshold), norm_error=int(solver._norm_error),
      inner_iterations=int(solver.inner_iterations), min_iterations=int(solver.min_iterations),
      max_iterations=int(solver.max_iterations), momentum=float(solver.momentum.value),
      tile_config=int(tile_config), use_graph=1 if solver.use_graph else 0, init_mode=_lib.INIT_ZEROS)
  lib = _lib.load_library()
  nbytes = lib.ottx_sinkhorn_batched_workspace_bytes(ctypes.byref(bt), ctypes.byref(opts))
  if nbytes == 0:
    raise ValueError("invalid batched problem")
  ws = _lib.workspace(nbytes, device)
  # zeros, -inf at zero weights (initializers.py:95-98), set by the solve itself
  f = torch.empty(xs.shape[0], dtype=torch.float32, device=device)
  g = torch.empty(ys.shape[0], dtype=torch.float32, device=device)
  errors = torch.empty((K, solver.outer_iterations), dtype=torch.float32, device=device)
  reg = (ctypes.c_double * K)()
  eps_out = (ctypes.c_double * K)()
  conv = (ctypes.c_int32 * K)()
  iters = (ctypes.c_int32 * K)()
  out = _lib.BatchOut(f=f.data_ptr(), g=g.data_ptr(), errors=errors.data_ptr(), reg_ot_cost=reg,
                      epsilon=eps_out, converged=conv, n_iters=iters)
  _lib.check("ottx_sinkhorn_batched_solve", lib.ottx_sinkhorn_batched_solve(
      ctypes.byref(bt), ctypes.byref(opts), ctypes.byref(out), ws.data_ptr(), ws.numel(), _lib.stream_ptr()))
  del ws
  # the host arrays as tensors without per-element conversions
  return SegmentSinkhornOutput(
      f=f, g=g, errors=errors, reg_ot_cost=torch.from_numpy(np.ctypeslib.as_array(reg).copy()),
      converged=torch.from_numpy(np.ctypeslib.as_array(conv) != 0),
      n_iters=torch.from_numpy(np.ctypeslib.as_array(iters).astype(np.int64)),
      epsilon=torch.from_numpy(np.ctypeslib.as_array(eps_out).copy()), x_offsets=x_off, y_offsets=y_off, a=a,
      b=b, n_blocks=int(out.n_blocks))


def segment_sinkhorn(
    x, y, num_segments: Optional[int] = None, max_measure_size: Optional[int] = None,
    cost_fn: Optional[costs.CostFn] = None, segment_ids_x=None, segment_ids_y=None,
    indices_are_sorted: bool = False, num_per_segment_x: Optional[Tuple[int, ...]] = None,
    num_per_segment_y: Optional[Tuple[int, ...]] = None, weights_x=None, weights_y=None,
    sinkhorn_kwargs: Mapping[str, Any] = MappingProxyType({}), **kwargs: Any) -> torch.Tensor:
  """Regularized OT cost of every segment pair (tools/segment_sinkhorn.py:24-138).

  Same arguments as OTT; ``kwargs`` go to the point cloud (``epsilon``,
  ``relative_epsilon``, ``scale_cost`` and, beyond OTT, ``precision``).
  Returns a float64 tensor [num_segments] on the CPU.
  """
  out = segment_sinkhorn_outputs(
      x, y, num_segments=num_segments, max_measure_size=max_measure_size, cost_fn=cost_fn,
      segment_ids_x=segment_ids_x, segment_ids_y=segment_ids_y, indices_are_sorted=indices_are_sorted,
      num_per_segment_x=num_per_segment_x, num_per_segment_y=num_per_segment_y, weights_x=weights_x,
      weights_y=weights_y, sinkhorn_kwargs=sinkhorn_kwargs, **kwargs)
  return out.reg_ot_cost
