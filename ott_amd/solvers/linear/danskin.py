"""A differentiable ``reg_ot_cost`` for PyTorch autograd.

The reference differentiates ``linear.solve(geom, a, b).reg_ot_cost`` with
jax.grad; under the default ``use_danskin`` its f and g are
``stop_gradient``'ed (sinkhorn.py:329-335), so the gradient is that of
compute_kl_reg_cost (sinkhorn.py:224-277) at fixed potentials.
:func:`reg_ot_cost` returns the solve's reg_ot_cost as a 0-dim float64 tensor
whose backward gives those gradients for x, y, a and b
(``SinkhornOutput.reg_ot_cost_grad``).
"""

from __future__ import annotations

from typing import Any, Optional

import torch

from ott_amd.geometry.pointcloud import PointCloud
from ott_amd.solvers.linear._solve import solve

__all__ = ["reg_ot_cost"]


class _RegOTCost(torch.autograd.Function):

  @staticmethod
  def forward(ctx, x, y, a, b, geom_kwargs, solve_kwargs):
    geom = PointCloud(x.detach(), y.detach(), **geom_kwargs)
    out = solve(geom, a=None if a is None else a.detach(), b=None if b is None else b.detach(), **solve_kwargs)
    ctx.out = out
    ctx.has_a, ctx.has_b = a is not None, b is not None
    # gradients go back to each input's own device
    ctx.devices = tuple(None if t is None else t.device for t in (x, y, a, b))
    return torch.tensor(out.reg_ot_cost, dtype=torch.float64, device=x.device)

  @staticmethod
  def backward(ctx, grad_out):
    out = ctx.out
    w = grad_out.to(device=out.f.device, dtype=torch.float64)
    need = ctx.needs_input_grad
    gx = gy = ga = gb = None
    if need[0] or need[1]:
      geom = out.geom
      if need[0]:
        gx = (w * geom.cost_gradient(out.f, out.g, side=0).double()).float()
      if need[1]:
        gy = (w * geom.cost_gradient(out.f, out.g, side=1).double()).float()
    if (need[2] and ctx.has_a) or (need[3] and ctx.has_b):
      wa, wb = out._weight_grads()
      ga = (w * wa).float() if need[2] and ctx.has_a else None
      gb = (w * wb).float() if need[3] and ctx.has_b else None
    grads = [None if gr is None else gr.to(dev) for gr, dev in zip((gx, gy, ga, gb), ctx.devices)]
    return (*grads, None, None)


def reg_ot_cost(x: torch.Tensor, y: torch.Tensor, a: Optional[torch.Tensor] = None,
                b: Optional[torch.Tensor] = None, *, epsilon: float, scale_cost: float = 1.0,
                precision: str = "bf16_split", **solve_kwargs: Any) -> torch.Tensor:
  """``linear.solve(PointCloud(x, y, epsilon, scale_cost), a, b, **solve_kwargs).reg_ot_cost``
  as a differentiable 0-dim tensor (gradients at fixed potentials, as the
  reference's default ``use_danskin``).  Balanced problems with an absolute
  epsilon and a float ``scale_cost``; x and y of at most 64 features."""
  if not isinstance(epsilon, (int, float)) or not isinstance(scale_cost, (int, float)):
    raise NotImplementedError("reg_ot_cost gradients need an absolute epsilon and a float scale_cost")
  if float(solve_kwargs.get("tau_a", 1.0)) != 1.0 or float(solve_kwargs.get("tau_b", 1.0)) != 1.0:
    raise NotImplementedError("reg_ot_cost gradients support balanced problems only")
  if x.shape[1] > 64:
    raise NotImplementedError("reg_ot_cost gradients support at most 64 features")
  # the gradient here holds f and g fixed, which is the reference's gradient
  # only under use_danskin (the default, sinkhorn.py:329-335); without it the
  # reference differentiates through the potentials as well
  if solve_kwargs.get("use_danskin", True) is False or (
      "implicit_diff" in solve_kwargs and solve_kwargs["implicit_diff"] is None and
      solve_kwargs.get("use_danskin") is None):
    raise NotImplementedError("reg_ot_cost gradients are the use_danskin ones (f and g held fixed)")
  geom_kwargs = dict(epsilon=float(epsilon), scale_cost=float(scale_cost), precision=precision)
  return _RegOTCost.apply(x, y, a, b, geom_kwargs, solve_kwargs)
