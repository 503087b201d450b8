"""std(C) on the GPU and the epsilons built on it (geometry.py:119-171):
relative_epsilon="std" and the reference's default, epsilon=None with
relative_epsilon=None, i.e. 0.05 * std(C).

The GPU forms std(C) in closed form from per-cloud moments; the checks below
compute it from every entry of C instead (blocked, float64), which shares
nothing with the moment formula.
"""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu


def _std_brute(x, y, scale=1.0, block=2048):
  """sqrt(relu(mean(C^2) - mean(C)^2)) over all n m entries, C formed from
  the differences block by block in float64."""
  x = np.asarray(x, np.float64)
  y = np.asarray(y, np.float64)
  s1 = s2 = 0.0
  nx = (x * x).sum(1)
  ny = (y * y).sum(1)
  for i in range(0, x.shape[0], block):
    xb = x[i:i + block]
    c = scale * (nx[i:i + block, None] + ny[None, :] - 2.0 * xb @ y.T)
    s1 += c.sum()
    s2 += (c * c).sum()
  nm = x.shape[0] * y.shape[0]
  var = s2 / nm - (s1 / nm) ** 2
  return math.sqrt(max(var, 0.0)), s1 / nm


def _geom(x, y, **kw):
  from ott_amd.geometry import pointcloud
  return pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), **kw)


@pytest.mark.parametrize("n,m,d,offset", [(1000, 1200, 2, 0.0), (777, 333, 64, 3.0), (512, 2049, 128, 0.5),
                                          (3000, 10, 5, 10.0), (1, 1, 3, 0.0), (16384, 16384, 64, 0.0)])
def test_std_cost_matches_every_entry(n, m, d, offset):
  rng = np.random.default_rng(n + d)
  x = (rng.normal(size=(n, d)) / np.sqrt(d) + offset).astype(np.float32)
  y = (rng.normal(size=(m, d)) / np.sqrt(d) + offset + 0.3 / np.sqrt(d)).astype(np.float32)
  print(f"std(C) n={n} m={m} d={d} offset={offset}", flush=True)
  for scale in (1.0, 0.37):
    geom = _geom(x, y, scale_cost=1.0 / scale)
    want, want_mean = _std_brute(x, y, scale)
    got = geom.std_cost_matrix
    # sqrt of a float64 difference: near std = 0 its floor is ~1e-8 mean(C)
    assert math.isclose(got, want, rel_tol=1e-6, abs_tol=1e-7 * abs(want_mean)), (got, want)
    assert math.isclose(geom.mean_cost_matrix, want_mean, rel_tol=1e-9, abs_tol=1e-12)
  if n * m <= 4e6:
    assert math.isclose(got, oracle.std_cost(x, y, scale), rel_tol=1e-6, abs_tol=1e-7 * abs(want_mean))


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("mode", ["default", "std", "std_mult"])
def test_solve_with_std_epsilon(precision, mode):
  """linear.solve(PointCloud(x, y), a, b) with the reference's defaults and
  the "std" relative epsilon, resolved on the device, against the oracle."""
  from ott_amd.solvers import linear
  rng = np.random.default_rng(5)
  n, m, d = 700, 900, 16
  x = (rng.normal(size=(n, d)) / 4).astype(np.float32)
  y = (rng.normal(size=(m, d)) / 4 + 0.1).astype(np.float32)
  a = rng.uniform(0.2, 1.2, n)
  b = rng.uniform(0.2, 1.2, m)
  a = (a / a.sum()).astype(np.float32)
  b = (b / b.sum()).astype(np.float32)
  kw = {"default": dict(), "std": dict(relative_epsilon="std"),
        "std_mult": dict(epsilon=0.2, relative_epsilon="std")}[mode]
  geom = _geom(x, y, precision=precision, **kw)
  out = linear.solve(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=kw.get("epsilon"), relative_epsilon=kw.get("relative_epsilon"))
  want_eps = oracle.resolve_epsilon(x, y, kw.get("epsilon"), kw.get("relative_epsilon"))
  assert math.isclose(geom.epsilon, want_eps, rel_tol=1e-6)
  tol = {"fp32": 1e-4, "bf16_split": 1e-3}[precision]
  assert out.converged and ref.converged
  assert out.n_iters == ref.n_iters
  assert abs(out.reg_ot_cost - ref.reg_ot_cost) <= tol * abs(ref.reg_ot_cost)
  f = out.f.cpu().numpy()
  assert np.abs(f - ref.f).max() <= tol * max(1.0, np.abs(ref.f).max())


def test_sinkhorn_divergence_with_default_epsilon():
  """sinkdiv(x, y) with no epsilon: the reference shares the (x, y) geometry's
  default epsilon, 0.05 std C(x, y), with the symmetric terms."""
  from ott_amd.tools import sinkhorn_divergence as sd
  rng = np.random.default_rng(8)
  x = rng.normal(size=(400, 8)).astype(np.float32)
  y = (rng.normal(size=(300, 8)) + 0.5).astype(np.float32)
  out = sd.sinkdiv(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
  eps = oracle.resolve_epsilon(x, y, None, None)
  kw = dict(epsilon=eps)
  xy = oracle.solve_pointcloud(x, y, **kw).reg_ot_cost
  xx = oracle.solve_pointcloud(x, x, **kw).reg_ot_cost
  yy = oracle.solve_pointcloud(y, y, **kw).reg_ot_cost
  want = xy - 0.5 * (xx + yy)
  got = float(out[0]) if isinstance(out, tuple) else float(out.divergence)
  assert abs(got - want) <= 1e-4 * max(1.0, abs(xy)), (got, want)
