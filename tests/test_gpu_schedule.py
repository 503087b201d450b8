"""Epsilon schedules (geometry/epsilon_scheduler.py:24-60) on the GPU solve
against the FP64 oracle.

Iteration t runs at ``max(init decay^t, 1) eps``; the marginal error and the
outputs use eps itself (sinkhorn.py:879-938).  The cases cover a schedule that
reaches eps in the second block (inner 10: the first block's error comes from the
separate error pass, later ones from the fused g-update), one that reaches it
after many two-iteration blocks, and ``decay = 1`` (never reaches eps: every
error comes from the error pass and the solve runs to max_iterations), each
balanced, unbalanced on both sides (the row error then needs its own pass at
eps while the schedule is above it) and unbalanced on one side.
"""

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle
from tests.test_gpu_parity import _check_solve, _cloud

pytestmark = pytest.mark.gpu


def _solve_scheduled(x, y, a, b, precision, eps, init, decay, **kw):
  from ott_amd.geometry import epsilon_scheduler, pointcloud
  from ott_amd.solvers import linear
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                               epsilon=epsilon_scheduler.Epsilon(eps, init=init, decay=decay),
                               precision=precision)
  return linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), **kw)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("inner,init,decay,max_it", [(10, 10.0, 0.8, 2000), (2, 30.0, 0.9, 2000),
                                                     (5, 2.0, 1.0, 60)])
@pytest.mark.parametrize("taus", [(1.0, 1.0), (0.9, 0.8), (1.0, 0.85)])
def test_epsilon_schedule_matches_oracle(precision, inner, init, decay, max_it, taus):
  x, y, a, b = _cloud(300, 257, 6, seed=12, zero_a=[4])
  eps = 0.02
  kw = dict(threshold=1e-4, inner_iterations=inner, max_iterations=max_it, tau_a=taus[0], tau_b=taus[1])
  out = _solve_scheduled(x, y, a, b, precision, eps, init, decay, **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, eps_init=init, eps_decay=decay, **kw)
  assert out.n_iters > inner
  if decay == 1.0:
    assert not ref.converged and ref.n_iters == max_it
  _check_solve(out, ref, precision)


def test_epsilon_schedule_is_refused_with_the_recentred_iteration():
  x, y, a, b = _cloud(64, 48, 3, seed=2)
  with pytest.raises(NotImplementedError):
    _solve_scheduled(x, y, a, b, "fp32", 0.05, 4.0, 0.5, tau_a=0.9, tau_b=0.9, recenter_potentials=True)
