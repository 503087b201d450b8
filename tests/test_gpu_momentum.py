"""Adaptive momentum, ``Momentum(start > 0)`` (acceleration.py:138-172), on the
GPU solve against the FP64 oracle.

From iteration ``start`` on, both updates use Lehmann's weight from the errors
of blocks start // inner - 2 and - 1 when the latter is below
``error_threshold``.  The cases cover ``start`` a multiple of
``inner_iterations`` (the block-ending g-update of iteration ``start`` is
blended after its block's error is final), a start inside a block, the graph
path (inner 10), and a threshold that holds and one that does not.
"""

import math

import numpy as np
import pytest

from oracle import sinkhorn64 as oracle
from tests.test_gpu_parity import _check_solve, _cloud, _solve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("inner,start,error_threshold", [(2, 6, math.inf), (3, 7, math.inf), (10, 30, math.inf),
                                                         (5, 10, 0.5), (5, 10, 0.05)])
def test_adaptive_momentum_matches_oracle(precision, inner, start, error_threshold):
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(300, 257, 6, seed=12, zero_a=[4])
  kw = dict(threshold=1e-4, inner_iterations=inner, max_iterations=2000)
  mom = acceleration.Momentum(start=start, error_threshold=error_threshold)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.02), momentum=mom, **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.02, momentum_start=start,
                                momentum_error_threshold=error_threshold, **kw)
  w = oracle.momentum_weight(ref.errors, start, 1.0, start, error_threshold, inner)
  assert (w > 1.0) == (error_threshold > 0.1)
  _check_solve(out, ref, precision)


def test_adaptive_momentum_limits():
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(64, 48, 3, seed=2)
  with pytest.raises(NotImplementedError):  # needs two recorded blocks before start
    _solve(x, y, a, b, "fp32", dict(epsilon=0.05), inner_iterations=5, momentum=acceleration.Momentum(start=5))


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("inner,start,init,decay", [(5, 10, 4.0, 0.9), (10, 30, 10.0, 0.8)])
def test_adaptive_momentum_with_epsilon_schedule(precision, inner, start, init, decay):
  """Both per-iteration options together: the schedule's error pass and the
  momentum blend after it run in the same blocks."""
  import torch
  from ott_amd.geometry import epsilon_scheduler, pointcloud
  from ott_amd.solvers import linear
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(300, 257, 6, seed=12, zero_a=[4])
  kw = dict(threshold=1e-4, inner_iterations=inner, max_iterations=2000)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                               epsilon=epsilon_scheduler.Epsilon(0.02, init=init, decay=decay), precision=precision)
  out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(),
                     momentum=acceleration.Momentum(start=start), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.02, momentum_start=start, eps_init=init, eps_decay=decay,
                                **kw)
  assert ref.converged
  _check_solve(out, ref, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("tau_a,tau_b,recenter", [(0.9, 0.8, False), (1.0, 0.85, False), (0.9, 0.8, True)])
def test_adaptive_momentum_unbalanced(precision, tau_a, tau_b, recenter):
  """Unbalanced problems: the block-ending f-update measures the row error
  that the weight reads beside the column error (sinkhorn.py:170-188), and
  the recentred iteration blends its own updates."""
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(300, 257, 6, seed=13, zero_b=[7])
  kw = dict(threshold=1e-4, inner_iterations=3, max_iterations=2000, tau_a=tau_a, tau_b=tau_b,
            recenter_potentials=recenter)
  mom = acceleration.Momentum(start=9)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.02), momentum=mom, **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.02, momentum_start=9, **kw)
  assert ref.converged
  _check_solve(out, ref, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("norm_error,taus", [(2, (1.0, 1.0)), (3, (1.0, 1.0)), (2, (0.9, 0.85))])
def test_adaptive_momentum_with_other_norms(precision, norm_error, taus):
  """Momentum(start) with norm_error != 1: the reference adds an L1 column to
  its errors (sinkhorn.py:1032-1038) yet writes the first norm's error into
  every column (one_iteration, :905-931), so Lehmann's weight reads the
  p-norm errors; the oracle records them the same way."""
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(300, 257, 6, seed=14, zero_a=[3])
  # epsilon and threshold chosen so that the threshold crossing is clear (the
  # last two block errors at about 1.6 and 0.7 of it)
  kw = dict(threshold=1e-5, inner_iterations=2, max_iterations=2000, norm_error=norm_error, tau_a=taus[0],
            tau_b=taus[1])
  out = _solve(x, y, a, b, precision, dict(epsilon=0.1), momentum=acceleration.Momentum(start=6), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, momentum_start=6, **kw)
  assert ref.converged
  w = oracle.momentum_weight(ref.errors, 6, 1.0, 6, math.inf, 2)
  assert w > 1.0
  _check_solve(out, ref, precision)
