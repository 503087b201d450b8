"""Solver outputs and hooks beyond the potentials: progress_fn calls
(sinkhorn.py:933-937), ent_reg_cost / dual_cost (sinkhorn.py:337-370) and
adaptive momentum on unbalanced problems, against the oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu


def _data(seed, n=300, m=257, d=8):
  rng = np.random.default_rng(seed)
  x = (rng.normal(size=(n, d)) / 3).astype(np.float32)
  y = (rng.normal(size=(m, d)) / 3 + 0.1).astype(np.float32)
  a = rng.uniform(0.2, 1.2, n)
  b = rng.uniform(0.2, 1.2, m)
  return x, y, (a / a.sum()).astype(np.float32), (b / b.sum()).astype(np.float32)


def _geom(x, y, **kw):
  from ott_amd.geometry import pointcloud
  return pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), **kw)


@pytest.mark.parametrize("kw", [dict(inner_iterations=5), dict(inner_iterations=3, min_iterations=7),
                                dict(inner_iterations=4, max_iterations=8, threshold=1e-12)])
def test_progress_fn_sees_the_reference_sequence(kw):
  from ott_amd.solvers import linear
  x, y, a, b = _data(1)
  got = []
  out = linear.solve(_geom(x, y, epsilon=0.05), torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                     progress_fn=lambda s: got.append((s[0], s[1], s[2], np.array(s[3].errors).ravel())), **kw)
  want = []
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05,
                                progress_fn=lambda s: want.append((s[0], s[1], s[2], np.array(s[3]).ravel())),
                                **kw)
  assert out.n_iters == ref.n_iters and len(got) == len(want) == ref.n_iters
  for (it, inner, mx, e), (it2, inner2, mx2, e2) in zip(got, want):
    assert (it, inner, mx) == (it2, inner2, mx2)
    assert np.array_equal(e == -1, e2 == -1) and np.array_equal(np.isinf(e), np.isinf(e2))
    fin = np.isfinite(e2) & (e2 != -1)
    np.testing.assert_allclose(e[fin], e2[fin], rtol=2e-3)


def test_progress_fn_exception_reaches_the_caller():
  from ott_amd.solvers import linear
  x, y, a, b = _data(2)

  def boom(status):
    raise KeyError("stop")
  with pytest.raises(KeyError):
    linear.solve(_geom(x, y, epsilon=0.05), progress_fn=boom)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_ent_reg_cost_and_dual_cost(precision):
  from ott_amd.solvers import linear
  x, y, a, b = _data(3)
  a[4] = 0.0
  a = (a / a.sum()).astype(np.float32)
  out = linear.solve(_geom(x, y, epsilon=0.05, precision=precision), torch.from_numpy(a).cuda(),
                     torch.from_numpy(b).cuda(), threshold=1e-6)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, threshold=1e-6)
  tol = {"fp32": 1e-4, "bf16_split": 1e-3}[precision]
  want_ent = oracle.ent_reg_cost(ref.reg_ot_cost, a.astype(np.float64), b.astype(np.float64), 0.05)
  want_dual = oracle.dual_cost(ref.f, ref.g, a.astype(np.float64), b.astype(np.float64))
  assert abs(out.ent_reg_cost - want_ent) <= tol * abs(want_ent)
  assert abs(out.dual_cost - want_dual) <= tol * max(1.0, abs(want_dual))


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_adaptive_momentum_unbalanced(precision):
  """Momentum(start) with tau < 1: the f-update's row error feeds the block's
  error, which sets Lehmann's weight (acceleration.py:138-160)."""
  from ott_amd.solvers import linear
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _data(4)
  # a threshold well above the float32 floor of the errors (about 1e-6 here)
  kw = dict(inner_iterations=2, max_iterations=60, threshold=1e-4)
  out = linear.solve(_geom(x, y, epsilon=0.02, precision=precision), torch.from_numpy(a).cuda(),
                     torch.from_numpy(b).cuda(), tau_a=0.9, tau_b=0.95,
                     momentum=acceleration.Momentum(start=6), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.02, tau_a=0.9, tau_b=0.95, momentum_start=6, **kw)
  tol = {"fp32": 1e-4, "bf16_split": 1e-3}[precision]
  same = (out.converged, out.n_iters) == (ref.converged, ref.n_iters)
  assert same, ((out.converged, out.n_iters), (ref.converged, ref.n_iters))
  rel = abs(out.reg_ot_cost - ref.reg_ot_cost) / abs(ref.reg_ot_cost)
  assert rel <= tol, rel
  err = float(np.abs(out.f.cpu().numpy() - ref.f).max())
  assert err <= tol * max(1.0, float(np.abs(ref.f).max())), err
