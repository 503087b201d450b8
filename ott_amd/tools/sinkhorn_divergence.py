"""Sinkhorn divergence (``ott.tools.sinkhorn_divergence``) on the ottx solver.

Mirrors tools/sinkhorn_divergence.py: ``sinkdiv`` (:135-176),
``sinkhorn_divergence`` (:178-250) and ``_sinkhorn_divergence`` (:253-351),
with the output type ``SinkhornDivergenceOutput`` (:35-131).  The three solves
run on the GPU:
- (x, y) with the caller's solver options;
- (x, x) and (y, y) with ``parallel_dual_updates=True`` and momentum 0.5, as the
  reference sets them for the symmetric terms (:298-318).  The solver runs such
  a symmetric problem with one pass per iteration, since f = g throughout.

The divergence is
``reg_ot_cost(xy) - (reg_ot_cost(xx) + reg_ot_cost(yy)) / 2 + eps (sum a - sum b)^2 / 2``
with eps the (x, y) geometry's epsilon (:320-324).  Low-rank solvers
(``rank > 0``), ``to_dual_potentials`` and ``segment_sinkhorn_divergence`` are
not provided.
"""

from __future__ import annotations

from types import MappingProxyType
from typing import Any, Mapping, NamedTuple, Optional, Tuple, Type

import torch

from ott_amd.geometry.pointcloud import PointCloud, as_device_f32
from ott_amd.solvers.linear import acceleration
from ott_amd.solvers.linear._solve import solve

__all__ = ["sinkdiv", "sinkhorn_divergence", "SinkhornDivergenceOutput"]

Potentials = Tuple[torch.Tensor, torch.Tensor]


class SinkhornDivergenceOutput(NamedTuple):
  """sinkhorn_divergence.py:35-131 (``factors`` is always ``None`` here)."""

  divergence: float
  geoms: Tuple[PointCloud, PointCloud, Optional[PointCloud]]
  a: torch.Tensor
  b: torch.Tensor
  potentials: Tuple[Potentials, Potentials, Potentials]
  factors: None
  errors: Tuple[Optional[torch.Tensor], Optional[torch.Tensor], Optional[torch.Tensor]]
  converged: Tuple[bool, bool, Optional[bool]]
  n_iters: Tuple[int, int, int]

  @property
  def is_low_rank(self) -> bool:
    return False


def sinkdiv(x, y, *, cost_fn=None, epsilon: Optional[float] = None,
            **kwargs: Any) -> Tuple[float, SinkhornDivergenceOutput]:
  """sinkhorn_divergence.py:135-176: the divergence between two point clouds."""
  return sinkhorn_divergence(PointCloud, x=x, y=y, cost_fn=cost_fn, epsilon=epsilon, **kwargs)


def sinkhorn_divergence(
    geom: Type[PointCloud],
    *args: Any,
    a=None,
    b=None,
    solve_kwargs: Mapping[str, Any] = MappingProxyType({}),
    static_b: bool = False,
    offset_static_b: Optional[float] = None,
    share_epsilon: bool = True,
    symmetric_sinkhorn: bool = True,
    **kwargs: Any,
) -> Tuple[float, SinkhornDivergenceOutput]:
  """sinkhorn_divergence.py:178-250.

  ``geom`` must be :class:`PointCloud` (or a subclass); ``args`` / ``kwargs``
  go to its ``prepare_divergences`` (x, y and the geometry's arguments).
  """
  if not (isinstance(geom, type) and issubclass(geom, PointCloud)):
    raise NotImplementedError("the Sinkhorn divergence is provided for PointCloud geometries only")
  if int(dict(solve_kwargs).get("rank", -1)) > 0:
    raise NotImplementedError("low-rank Sinkhorn (rank > 0) is outside ott_amd")
  geoms = geom.prepare_divergences(*args, static_b=static_b, **kwargs)
  geom_xy, geom_x, geom_y = (tuple(geoms) + (None,) * 3)[:3]
  num_a, num_b = geom_xy.shape
  device = geom_xy.device
  a = torch.full((num_a,), 1.0 / num_a, device=device) if a is None else as_device_f32(a, device)
  b = torch.full((num_b,), 1.0 / num_b, device=device) if b is None else as_device_f32(b, device)
  out = _sinkhorn_divergence(geom_xy, geom_x, geom_y, a=a, b=b, share_epsilon=share_epsilon,
                             symmetric_sinkhorn=symmetric_sinkhorn, offset_yy=offset_static_b,
                             **dict(solve_kwargs))
  return out.divergence, out


def _sinkhorn_divergence(geometry_xy: PointCloud, geometry_xx: PointCloud, geometry_yy: Optional[PointCloud],
                         a: torch.Tensor, b: torch.Tensor, share_epsilon: bool, symmetric_sinkhorn: bool,
                         offset_yy: Optional[float], **kwargs: Any) -> SinkhornDivergenceOutput:
  """sinkhorn_divergence.py:253-351 (Sinkhorn solver only).

  With ``share_epsilon`` the symmetric terms use the (x, y) geometry's
  epsilon (:235-240); it is resolved by the (x, y) solve, which therefore runs
  first, so a relative epsilon needs no separate pass over the clouds.
  """
  kwargs_symmetric = dict(kwargs)
  if symmetric_sinkhorn:
    # :298-318: parallel updates with momentum 0.5 for the symmetric terms
    kwargs_symmetric.update(parallel_dual_updates=True,
                            momentum=acceleration.Momentum(start=0, value=0.5), anderson=None)
  # Uniform weights are passed to the solves as None (the library sets
  # 1 / n itself); the symmetric solves need the same weights on both sides.
  a_solve = None if _is_uniform(a) else a
  b_solve = None if _is_uniform(b) else b
  out_xy = solve(geometry_xy, a=a_solve, b=b_solve, **kwargs)
  if share_epsilon:
    geometry_xx = geometry_xx.copy_epsilon(geometry_xy)
    if geometry_yy is not None:
      geometry_yy = geometry_yy.copy_epsilon(geometry_xy)
  out_xx = solve(geometry_xx, a=a_solve, b=a_solve, **kwargs_symmetric)
  if geometry_yy is None:
    # static_b: the (y, y) term is the given offset, or 0, as the reference's
    # empty output (:486-510) holds it
    reg_yy = 0.0 if offset_yy is None else float(offset_yy)
    pot_yy, conv_yy, it_yy = (None, None), None, 0
    err_yy = torch.tensor([-float("inf")], device=a.device)
  else:
    out_yy = solve(geometry_yy, a=b_solve, b=b_solve, **kwargs_symmetric)
    reg_yy, pot_yy = out_yy.reg_ot_cost, (out_yy.f, out_yy.g)
    err_yy, conv_yy, it_yy = out_yy.errors, out_yy.converged, out_yy.n_iters
  eps = geometry_xy.epsilon
  mass = float(a.double().sum() - b.double().sum())
  div = out_xy.reg_ot_cost - 0.5 * (out_xx.reg_ot_cost + reg_yy) + 0.5 * eps * mass ** 2
  return SinkhornDivergenceOutput(
      divergence=float(div),
      geoms=(geometry_xy, geometry_xx, geometry_yy),
      a=a,
      b=b,
      potentials=((out_xy.f, out_xy.g), (out_xx.f, out_xx.g), pot_yy),
      factors=None,
      errors=(out_xy.errors, out_xx.errors, err_yy),
      converged=(bool(out_xy.converged), bool(out_xx.converged), conv_yy),
      n_iters=(out_xy.n_iters, out_xx.n_iters, it_yy),
  )


def _is_uniform(w: torch.Tensor) -> bool:
  """Exactly 1 / n everywhere (the weights the library sets for None)."""
  n = w.shape[0]
  return bool(torch.all(w == torch.tensor(1.0 / n, dtype=w.dtype, device=w.device)).item())
