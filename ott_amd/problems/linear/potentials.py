"""Dual potential functions of a solved point-cloud problem.

Mirrors ``ott.problems.linear.potentials.DualPotentials`` (potentials.py:37-113)
as ``SinkhornOutput.to_dual_potentials`` (solvers/linear/sinkhorn.py:480-499)
builds it: ``f`` and ``g`` are soft c-transforms of the solved vectors
(``LinearProblem.potential_fn_from_dual_vec`` / ``_c_transform``,
problems/linear/linear_problem.py:111-171),

  f(x) = -eps log sum_j b_j exp((g_j - c(x, y_j)) / eps),
  g(y) = -eps log sum_i exp((f_i - c(x_i, y)) / eps) + eps log k,

the second as the reference evaluates it: its problem over the k new points
keeps uniform weights 1/k on them, which ``logsumexp(z, b=self.b, axis=0)``
broadcasts over the rows (so a single point gets no weight term at all).
Both are evaluated on the GPU as the log-sum-exp operator of a PointCloud between the
new points and the solved cloud (the reference builds that cloud without the
solved geometry's scale_cost, and so does this module).  ``transport`` is the
Gangbo-McCann map of the squared-Euclidean cost, x - grad f(x) / 2
(potentials.py:52-84 with costs.py:220-226 and h*(z) = |z|^2 / 4), which for
this f is the barycentric projection sum_j pi_j(x) y_j, pi_j(x) proportional
to b_j exp((g_j - |x - y_j|^2) / eps): one fused transport-apply launch.
"""

from __future__ import annotations

import math
from typing import Callable, Optional

import torch

from ott_amd.geometry.pointcloud import PointCloud, as_device_f32

__all__ = ["DualPotentials"]


class _CTransform:
  """x -> -eps log sum_k w_k exp((h_k - c(x, z_k)) / eps) over a solved cloud z
  (w = 1 when ``weights`` is None)."""

  def __init__(self, support: torch.Tensor, h: torch.Tensor, weights: Optional[torch.Tensor], eps: float,
               side: int, precision: str, cost_fn=None):
    self.support = support
    self.cost_fn = cost_fn
    self.eps = float(eps)
    self.side = side  # 1: f (x against the y cloud), 0: g (y against the x cloud)
    self.precision = precision
    with torch.no_grad():
      if weights is None:
        self.h_w = h.to(torch.float32).contiguous()
      else:
        w = weights.to(torch.float64)
        lw = torch.where(w > 0, torch.log(w), torch.full_like(w, -float("inf")))
        self.h_w = (h.to(torch.float64) + self.eps * lw).to(torch.float32)

  def _geom(self, pts: torch.Tensor) -> PointCloud:
    if self.side == 1:
      return PointCloud(pts, self.support, cost_fn=self.cost_fn, epsilon=self.eps, precision=self.precision)
    return PointCloud(self.support, pts, cost_fn=self.cost_fn, epsilon=self.eps, precision=self.precision)

  def values(self, pts: torch.Tensor):
    geom = self._geom(pts)
    zeros = torch.zeros((pts.shape[0],), dtype=torch.float32, device=pts.device)
    if self.side == 1:
      lse, _ = geom.apply_lse_kernel(zeros, self.h_w, self.eps, axis=1)
    else:
      lse, _ = geom.apply_lse_kernel(self.h_w, zeros, self.eps, axis=0)
    return -lse, geom

  def __call__(self, x) -> torch.Tensor:
    pts = as_device_f32(x, self.support.device)
    single = pts.ndim == 1
    pts = pts[None] if single else pts
    vals, _ = self.values(pts.contiguous())
    if self.side == 0 and pts.shape[0] > 1:
      vals = vals + self.eps * math.log(pts.shape[0])  # the uniform weights 1/k of the new points
    return vals[0] if single else vals

  def barycenter(self, x) -> torch.Tensor:
    """sum_k pi_k(x) z_k with pi_k(x) proportional to w_k exp((h_k - c(x, z_k)) / eps)."""
    pts = as_device_f32(x, self.support.device)
    single = pts.ndim == 1
    pts = (pts[None] if single else pts).contiguous()
    vals, geom = self.values(pts)
    coords = self.support.t().contiguous()
    if self.side == 1:
      # rows are the new points: P_ij = exp((vals_i + h_w_j - C_ij) / eps) sums to 1 over j
      out = geom.apply_transport_from_potentials(vals, self.h_w, coords, axis=1)
    else:
      out = geom.apply_transport_from_potentials(self.h_w, vals, coords, axis=0)
    out = out.t().contiguous()
    return out[0] if single else out


class _Debiased:
  """x -> t1(x) - t2(x) of two c-transforms over the same new points: the
  debiased potentials of a Sinkhorn divergence (sinkhorn_divergence.py:
  76-114).  With the squared-Euclidean cost grad t(x) = 2 (x - bary_t(x)), so
  the map x - grad(t1 - t2)(x) / 2 is bary_1(x) - bary_2(x) + x."""

  def __init__(self, t1: _CTransform, t2: _CTransform):
    self.t1 = t1
    self.t2 = t2
    self.support = t1.support

  def values(self, pts: torch.Tensor):
    v1, geom = self.t1.values(pts)
    v2, _ = self.t2.values(pts)
    return v1 - v2, geom

  def __call__(self, x) -> torch.Tensor:
    return self.t1(x) - self.t2(x)

  def barycenter(self, x) -> torch.Tensor:
    pts = as_device_f32(x, self.support.device)
    return self.t1.barycenter(pts) - self.t2.barycenter(pts) + pts


class DualPotentials:
  """potentials.py:37-113: the dual potential functions f and g of a solved
  problem, with ``transport`` and ``distance``."""

  def __init__(self, f: Optional[Callable], g: Optional[Callable], cost_fn=None):
    self.f = f
    self.g = g
    self.cost_fn = cost_fn

  def transport(self, vec, forward: bool = True) -> torch.Tensor:
    """potentials.py:52-84: x - grad f(x) / 2 (forward) or y - grad g(y) / 2."""
    fn = self.f if forward else self.g
    if not isinstance(fn, (_CTransform, _Debiased)):
      raise NotImplementedError("transport needs the c-transform potentials of to_dual_potentials")
    return fn.barycenter(vec)

  def distance(self, src, tgt) -> float:
    """potentials.py:86-101: mean f(src) + mean g(tgt), each point evaluated on
    its own (the reference vmaps the functions over the points)."""
    f, _ = self.f.values(torch.atleast_2d(as_device_f32(src, self.f.support.device)).contiguous())
    g, _ = self.g.values(torch.atleast_2d(as_device_f32(tgt, self.g.support.device)).contiguous())
    return float(f.double().mean() + g.double().mean())


def from_solution(geom: PointCloud, f: torch.Tensor, g: torch.Tensor, a: torch.Tensor, b: torch.Tensor,
                  epsilon: Optional[float] = None) -> DualPotentials:
  """SinkhornOutput.to_dual_potentials (sinkhorn.py:480-499)."""
  eps = geom.epsilon if epsilon is None else float(epsilon)
  f_fn = _CTransform(geom.y, g, b, eps, side=1, precision=geom.precision, cost_fn=geom.cost_fn)
  del a  # the reference's g transform takes no weights on the solved rows
  g_fn = _CTransform(geom.x, f, None, eps, side=0, precision=geom.precision, cost_fn=geom.cost_fn)
  return DualPotentials(f_fn, g_fn, cost_fn=geom.cost_fn)
