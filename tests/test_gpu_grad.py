"""Gradients of reg_ot_cost at fixed potentials (the reference's default
``use_danskin``, sinkhorn.py:329-335 with compute_kl_reg_cost :224-277)
against the oracle: ``SinkhornOutput.reg_ot_cost_grad`` at the potentials the
GPU solve returned (the point gradients run on ``ottx_cost_gradient``), and
the autograd path ``danskin.reg_ot_cost`` end to end against the oracle's
gradient at its own converged potentials."""

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle
from tests.test_gpu_parity import _cloud

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("n,m,d,scale", [(300, 257, 8, 1.0), (200, 131, 37, 0.5), (500, 500, 64, 1.0)])
def test_reg_ot_cost_grad_at_solve_potentials(precision, n, m, d, scale):
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  x, y, a, b = _cloud(n, m, d, seed=n + d, zero_a=[3])
  x = x + 1.0  # off-centre clouds: the per-pair differences must not cancel
  y = y + 1.0
  eps = 0.05
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=eps,
                               scale_cost=1.0 / scale, precision=precision)
  out = sinkhorn.Sinkhorn(threshold=1e-4)(linear_problem.LinearProblem(
      geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()))
  assert out.converged
  f = out.f.double().cpu().numpy()
  g = out.g.double().cpu().numpy()
  ref = oracle.reg_ot_cost_grad(x, y, a, b, f, g, eps, scale)
  got = [t.double().cpu().numpy() for t in out.reg_ot_cost_grad()]
  for name, gv, rv in zip("xyab", got, ref):
    assert gv.shape == rv.shape, name
    np.testing.assert_allclose(gv, rv, rtol=0, atol=2e-5 * float(np.abs(rv).max()) + 1e-12, err_msg=name)
  # zero weight: no coupling on that row; only the eps sum(b) term in a's gradient
  assert np.all(got[0][3] == 0.0) and got[2][3] == pytest.approx(eps * float(b.astype(np.float64).sum()))


def test_autograd_reg_ot_cost_matches_oracle_gradient():
  from ott_amd.solvers.linear import danskin
  x, y, a, b = _cloud(256, 200, 16, seed=5)
  eps = 0.05
  tx = torch.from_numpy(x).cuda().requires_grad_(True)
  ty = torch.from_numpy(y).cuda().requires_grad_(True)
  ta = torch.from_numpy(a).cuda().requires_grad_(True)
  tb = torch.from_numpy(b).cuda()
  cost = danskin.reg_ot_cost(tx, ty, ta, tb, epsilon=eps, precision="fp32", threshold=1e-6, max_iterations=5000)
  (2.0 * cost).backward()
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, threshold=1e-9, max_iterations=20000)
  assert ref.converged
  assert abs(float(cost) - ref.reg_ot_cost) <= 1e-4 * abs(ref.reg_ot_cost)
  gx, gy, ga, _ = oracle.reg_ot_cost_grad(x, y, a, b, ref.f, ref.g, eps)
  np.testing.assert_allclose(tx.grad.double().cpu().numpy(), 2.0 * gx, rtol=0, atol=1e-3 * np.abs(2 * gx).max())
  np.testing.assert_allclose(ty.grad.double().cpu().numpy(), 2.0 * gy, rtol=0, atol=1e-3 * np.abs(2 * gy).max())
  # weights: only differences are fixed by the balanced problem's gauge (f + c, g - c)
  dga = ta.grad.double().cpu().numpy()
  np.testing.assert_allclose(dga - dga.mean(), 2.0 * (ga - ga.mean()), rtol=0, atol=1e-3 * np.abs(2 * ga).max())
  assert tb.grad is None
