"""Parallel dual updates on symmetric problems and the Sinkhorn divergence
(tools/sinkhorn_divergence.py) against the FP64 oracle.

The symmetric terms of a divergence run ``parallel_dual_updates=True`` with
momentum 0.5 (sinkhorn_divergence.py:298-318); on x against x with a = b the
solver keeps f = g and does one pass per iteration.  Checked: iterations,
errors (row + column marginal, sinkhorn.py:160-189), potentials and
reg_ot_cost against the oracle's parallel solve; the divergence against three
oracle solves; zero divergence of a measure with itself; ``static_b``.
"""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle
from tests.test_gpu_parity import _check_solve

pytestmark = pytest.mark.gpu


def _points(n, d, seed, shift=0.0):
  rng = np.random.default_rng(seed)
  x = (rng.normal(size=(n, d)) / np.sqrt(d) + shift).astype(np.float32)
  w = rng.uniform(0.2, 1.2, size=n)
  return x, (w / w.sum()).astype(np.float32)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("inner", [1, 10])
def test_symmetric_parallel_solve_matches_oracle(precision, inner):
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import acceleration, sinkhorn
  x, a = _points(257, 8, seed=21)
  a[5] = 0.0  # a zero weight: -inf in both potentials
  a = (a / a.sum()).astype(np.float32)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), epsilon=0.05, relative_epsilon="mean",
                               precision=precision)
  ta = torch.from_numpy(a).cuda()
  solver = sinkhorn.Sinkhorn(threshold=1e-4, inner_iterations=inner, parallel_dual_updates=True,
                             momentum=acceleration.Momentum(start=0, value=0.5))
  out = solver(linear_problem.LinearProblem(geom, ta, ta))
  ref = oracle.solve_pointcloud(x, x, a, a, epsilon=0.05, relative_epsilon="mean", threshold=1e-4,
                                inner_iterations=inner, parallel_dual_updates=True, momentum=0.5)
  assert ref.converged and out.converged
  assert torch.equal(out.f, out.g)
  assert math.isinf(float(out.f[5])) and float(out.f[5]) < 0
  _check_solve(out, ref, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("case", ["xy", "xx_weights", "unbalanced", "warm"])
@pytest.mark.parametrize("inner,momentum", [(1, 0.5), (3, 0.5), (10, 0.7)])
def test_parallel_updates_general_matches_oracle(precision, case, inner, momentum):
  """parallel_dual_updates=True off the symmetric special case (sinkhorn.py:
  840-846): both updates from the previous state, errors as the row plus the
  column marginal error (sinkhorn.py:160-189), against the oracle's parallel
  iteration; x against y, x against x with different weights, unbalanced,
  and a warm start.  Balanced problems need momentum < 1 here: without it
  the parallel iteration oscillates in the gauge and never meets the
  threshold (the reference sets 0.5 for the divergence's symmetric terms)."""
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  x, a = _points(257, 5, seed=1)
  y, b = _points(190, 5, seed=2)
  kw = dict(threshold=1e-4, inner_iterations=inner, max_iterations=3000)
  taus = (0.9, 0.8) if case == "unbalanced" else (1.0, 1.0)
  if case == "xx_weights":
    y, b = x, np.full(257, 1.0 / 257, np.float32)
  tx = torch.from_numpy(x).cuda()
  ty = tx if case == "xx_weights" else torch.from_numpy(y).cuda()
  geom = pointcloud.PointCloud(tx, None if case == "xx_weights" else ty, epsilon=0.1, precision=precision)
  prob = linear_problem.LinearProblem(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                                      tau_a=taus[0], tau_b=taus[1])
  from ott_amd.solvers.linear import acceleration
  solver = sinkhorn.Sinkhorn(parallel_dual_updates=True, momentum=acceleration.Momentum(value=momentum), **kw)
  init = None
  if case == "warm":
    rng = np.random.default_rng(3)
    f0 = (rng.normal(size=257) * 0.01).astype(np.float32)
    g0 = (rng.normal(size=190) * 0.01).astype(np.float32)
    init = (torch.from_numpy(f0).cuda(), torch.from_numpy(g0).cuda())
  print(f"parallel {case} {precision} inner={inner} momentum={momentum}", flush=True)
  out = solver(prob, init=init)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, tau_a=taus[0], tau_b=taus[1], parallel_dual_updates=True,
                                momentum=momentum, init=None if init is None else (f0, g0), **kw)
  assert ref.converged
  _check_solve(out, ref, precision)


def _oracle_divergence(x, y, a, b, eps, **kw):
  xy = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, **kw)
  sym = dict(kw, parallel_dual_updates=True, momentum=0.5)
  xx = oracle.solve_pointcloud(x, x, a, a, epsilon=eps, **sym)
  yy = oracle.solve_pointcloud(y, y, b, b, epsilon=eps, **sym)
  div = xy.reg_ot_cost - 0.5 * (xx.reg_ot_cost + yy.reg_ot_cost) + 0.5 * eps * (a.sum() - b.sum()) ** 2
  return div, (xy, xx, yy)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_sinkhorn_divergence_matches_oracle(precision):
  from ott_amd.tools import sinkhorn_divergence as sd
  x, a = _points(300, 6, seed=3)
  y, b = _points(200, 6, seed=4, shift=0.3)
  kw = dict(threshold=1e-5, inner_iterations=1)
  div, out = sd.sinkdiv(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=0.05,
                        relative_epsilon="mean", precision=precision, a=a, b=b, solve_kwargs=kw)
  eps = oracle.resolve_epsilon(x, y, 0.05, "mean", 1.0)
  ref, outs = _oracle_divergence(x.astype(np.float64), y.astype(np.float64), a.astype(np.float64),
                                 b.astype(np.float64), eps, **kw)
  assert out.converged == (True, True, True)
  assert out.n_iters == tuple(o.n_iters for o in outs)
  assert math.isclose(out.geoms[1].epsilon, out.geoms[0].epsilon, rel_tol=1e-12)  # share_epsilon
  scale = max(abs(o.reg_ot_cost) for o in outs)
  assert abs(div - ref) <= 1e-5 * scale, (div, ref)
  assert div > 0.0
  for (f, g), o in zip(out.potentials, outs):
    np.testing.assert_allclose(f.cpu().numpy() - f.cpu().numpy()[0], o.f - o.f[0], atol=1e-3 * scale)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_sinkhorn_divergence_of_a_measure_with_itself_is_zero(precision):
  """x against x: the (x, y) solve (sequential) and the symmetric one (parallel)
  reach the same fixed point, so the divergence vanishes to solver accuracy."""
  from ott_amd.tools import sinkhorn_divergence as sd
  x, a = _points(256, 5, seed=7)
  tx = torch.from_numpy(x).cuda()
  div, out = sd.sinkdiv(tx, tx, epsilon=0.1, relative_epsilon="mean", precision=precision, a=a, b=a,
                        solve_kwargs=dict(threshold=1e-5, inner_iterations=1, max_iterations=5000))
  assert all(out.converged)
  scale = max(abs(float(out.potentials[0][0].double().abs().max())), 1.0)
  assert abs(div) <= 1e-4 * scale, div


def test_sinkhorn_divergence_static_b():
  """static_b: the (y, y) term is offset_static_b (or 0) and is not solved
  (sinkhorn_divergence.py:325-333, 486-510)."""
  from ott_amd.tools import sinkhorn_divergence as sd
  x, a = _points(128, 4, seed=9)
  y, b = _points(96, 4, seed=10, shift=0.2)
  tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
  kw = dict(epsilon=0.1, solve_kwargs=dict(threshold=1e-5, inner_iterations=1))
  full, out_full = sd.sinkdiv(tx, ty, **kw)
  static, out_static = sd.sinkdiv(tx, ty, static_b=True, **kw)
  assert out_static.geoms[2] is None and out_static.potentials[2] == (None, None)
  assert out_static.n_iters[2] == 0 and out_static.converged[2] is None
  # the yy term of the full divergence, recovered from both results
  reg_yy = 2.0 * (static - full)  # full = xy - (xx + yy)/2, static = xy - xx/2
  offset, out_off = sd.sinkdiv(tx, ty, static_b=True, offset_static_b=reg_yy, **kw)
  assert math.isclose(offset, full, rel_tol=1e-6, abs_tol=1e-9)
