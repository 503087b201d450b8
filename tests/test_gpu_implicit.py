"""Implicit differentiation of the potentials (implicit_differentiation.py:81-269,
sinkhorn.py:1102-1136) against central differences of FP64 oracle solves.

L = <gr_f, f> + <gr_g, g> is differentiated with respect to x, y, a and b along
random directions.  In the balanced case the potentials carry a free constant
(f + c, g - c), so the cotangents are chosen with sum gr_f = sum gr_g (L does
not see the constant) and the weight directions keep the masses equal.
"""

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu

EPS = 0.1


def _data(seed):
  rng = np.random.default_rng(seed)
  n, m, d = 60, 45, 3
  x = rng.normal(size=(n, d)) * 0.7
  y = rng.normal(size=(m, d)) * 0.7 + 0.3
  a = rng.uniform(0.5, 1.5, n)
  b = rng.uniform(0.5, 1.5, m)
  return rng, x, y, a / a.sum(), b / b.sum()


def _oracle_l(x, y, a, b, gr_f, gr_g, taus):
  out = oracle.solve_pointcloud(x, y, a, b, epsilon=EPS, tau_a=taus[0], tau_b=taus[1], threshold=1e-13,
                                inner_iterations=1, max_iterations=200000)
  assert out.converged
  return float(gr_f @ out.f + gr_g @ out.g)


@pytest.mark.parametrize("taus", [(1.0, 1.0), (0.8, 0.9)])
@pytest.mark.parametrize("wrt", ["x", "y", "a", "b"])
def test_implicit_vjp_matches_central_differences(taus, wrt):
  from ott_amd.solvers.linear import implicit_differentiation as imp
  rng, x, y, a, b = _data(5)
  n, m = len(a), len(b)
  gr_f = rng.normal(size=n)
  gr_g = rng.normal(size=m)
  if taus == (1.0, 1.0):
    gr_g += (gr_f.sum() - gr_g.sum()) / m
  xt = torch.tensor(x, dtype=torch.float32, device="cuda", requires_grad=True)
  yt = torch.tensor(y, dtype=torch.float32, device="cuda", requires_grad=True)
  at = torch.tensor(a, dtype=torch.float32, device="cuda", requires_grad=True)
  bt = torch.tensor(b, dtype=torch.float32, device="cuda", requires_grad=True)
  f, g = imp.potentials(xt, yt, at, bt, epsilon=EPS, precision="fp32", tau_a=taus[0], tau_b=taus[1],
                        threshold=1e-6, inner_iterations=1, max_iterations=20000)
  loss = (torch.from_numpy(gr_f).cuda() * f.double()).sum() + (torch.from_numpy(gr_g).cuda() * g.double()).sum()
  loss.backward()
  base = dict(x=x, y=y, a=a, b=b)
  grad = dict(x=xt.grad, y=yt.grad, a=at.grad, b=bt.grad)[wrt].double().cpu().numpy()
  direction = rng.normal(size=base[wrt].shape)
  if wrt in ("a", "b"):
    direction -= direction.mean()  # keeps the mass (and a balanced problem balanced)
  h = 1e-5
  plus = dict(base)
  minus = dict(base)
  plus[wrt] = base[wrt] + h * direction
  minus[wrt] = base[wrt] - h * direction
  fd = (_oracle_l(**plus, gr_f=gr_f, gr_g=gr_g, taus=taus) - _oracle_l(**minus, gr_f=gr_f, gr_g=gr_g, taus=taus)) / (2 * h)
  got = float((grad * direction).sum())
  assert abs(got - fd) <= 2e-3 * max(abs(fd), 1e-3), (got, fd)
