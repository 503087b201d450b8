"""The reference's numerical failure conventions on the GPU path, against the
oracle (sinkhorn.py:940-962 stop rule, :1013-1023 converged):

- a NaN input makes the block's error NaN: the loop stops after that block
  with converged False, the NaN row confined to its own costs (its potential
  is eps log a, geometry.py:376-398 maps a non-finite log-sum-exp to 0);
- a +inf error also stops the loop, and since the unreached rows of
  ``errors`` stay -1, converged = not any(isnan(errors)) and errors[-1] <
  threshold reads True: the reference's quirk, kept;
- a balanced problem with sum a != sum b never meets the threshold and runs
  to max_iterations, unconverged.
"""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16_split": 1e-3}


def _data(seed=0, n=60, m=50, d=3):
  rng = np.random.default_rng(seed)
  x = rng.normal(size=(n, d)).astype(np.float32)
  y = (rng.normal(size=(m, d)) + 0.2).astype(np.float32)
  a = np.full(n, 1.0 / n, dtype=np.float32)
  b = np.full(m, 1.0 / m, dtype=np.float32)
  return x, y, a, b


def _solve(x, y, a, b, precision, eps, init=None, **kw):
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=eps,
                               precision=precision)
  prob = linear_problem.LinearProblem(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
  if init is not None:
    init = tuple(torch.from_numpy(v.astype(np.float32)).cuda() for v in init)
  return sinkhorn.Sinkhorn(**kw)(prob, init=init)


def _same_layout(got, ref):
  g = got.errors.cpu().numpy().astype(np.float64)
  assert np.array_equal(np.isnan(g), np.isnan(ref.errors))
  assert np.array_equal(g == -1, ref.errors == -1)
  assert np.array_equal(np.isposinf(g), np.isposinf(ref.errors))
  assert got.converged == ref.converged
  assert got.n_iters == ref.n_iters


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("where", ["x", "y"])
def test_nan_input_stops_after_one_block(precision, where):
  x, y, a, b = _data()
  if where == "x":
    x[3, 1] = np.nan
  else:
    y[7, 0] = np.nan
  kw = dict(inner_iterations=5, max_iterations=50)
  print(f"nan in {where}, {precision}", flush=True)
  out = _solve(x, y, a, b, precision, 0.1, **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, **kw)
  assert not ref.converged and ref.n_iters == 5 and np.isnan(ref.errors[0])
  _same_layout(out, ref)
  assert math.isnan(out.reg_ot_cost) and math.isnan(ref.reg_ot_cost)
  f = out.f.cpu().numpy().astype(np.float64)
  g = out.g.cpu().numpy().astype(np.float64)
  assert np.array_equal(np.isfinite(f), np.isfinite(ref.f))
  assert np.array_equal(np.isfinite(g), np.isfinite(ref.g))
  tol = TOL[precision]
  np.testing.assert_allclose(f, ref.f, atol=tol * max(1.0, np.abs(ref.f).max()))
  np.testing.assert_allclose(g, ref.g, atol=tol * max(1.0, np.abs(ref.g).max()))


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_infinite_error_stops_and_reads_converged(precision):
  """Over-relaxation (momentum 1.9) from f0 = 1000 overflows the marginal:
  the first error is +inf, the loop stops, and converged is True."""
  x, y, a, b = _data(seed=1)
  init = (np.full(60, 1000.0), np.zeros(50))
  kw = dict(inner_iterations=1, max_iterations=20)
  print(f"inf error, {precision}", flush=True)
  from ott_amd.solvers.linear import acceleration
  out = _solve(x, y, a, b, precision, 0.01, init=init, momentum=acceleration.Momentum(value=1.9), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.01, momentum=1.9, init=init, **kw)
  assert ref.converged and ref.n_iters == 1 and np.isposinf(ref.errors[0])
  _same_layout(out, ref)
  assert out.reg_ot_cost == -math.inf and ref.reg_ot_cost == -math.inf


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_unequal_masses_run_to_max_iterations(precision):
  x, y, a, b = _data(seed=2)
  b = (b * 1.2).astype(np.float32)
  kw = dict(inner_iterations=5, max_iterations=50)
  print(f"unequal masses, {precision}", flush=True)
  out = _solve(x, y, a, b, precision, 0.1, **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, **kw)
  assert not ref.converged and ref.n_iters == 50
  _same_layout(out, ref)
  np.testing.assert_allclose(out.errors.cpu().numpy(), ref.errors, rtol=1e-3)
