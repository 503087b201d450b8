"""What the trained DeepSets warm start gains over zero initialisation.

Both shipped networks (ott_amd/initializers/neural/weights) are run as the
Sinkhorn initializer on held-out problems of the families they were trained
on; the converged solver judges the result.  Checked: fewer iterations to the
1e-3 marginal error than zero initialisation (DefaultInitializer,
initializers.py:70-131), convergence, and reg_ot_cost against the FP64 oracle
(rel 1e-3 on the BF16-split path).
"""

import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _solve(x, y, initializer, precision="bf16_split", eps_mult=0.05):
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  geom = pointcloud.PointCloud(x, y, epsilon=eps_mult, relative_epsilon="mean", precision=precision)
  prob = linear_problem.LinearProblem(geom)
  return sinkhorn.Sinkhorn(threshold=1e-3, inner_iterations=1, max_iterations=5000,
                           initializer=initializer)(prob)


def _initializer(d):
  from ott_amd.initializers.neural import deepsets
  path = deepsets.shipped_model_path(d)
  assert path is not None, f"no shipped weights for d={d}"
  return deepsets.DeepSetsInitializer(deepsets.DeepSets.load(path))


def test_gaussian_n16384_d64_one_iteration_fewer():
  """The benchmark workload (n=m=16384, d=64): 3 iterations from zeros, 2 from the network."""
  from ott_amd import datasets
  with open(os.path.join(HERE, "gaussian_shifted_n16384_d64.json")) as fh:
    gold = json.load(fh)
  pb = datasets.gaussian_shifted(16384, 16384, 64, seed=1)
  x, y = torch.from_numpy(pb.x).cuda(), torch.from_numpy(pb.y).cuda()
  zero = _solve(x, y, None)
  learned = _solve(x, y, _initializer(64))
  assert zero.converged and learned.converged
  assert learned.n_iters < zero.n_iters, (learned.n_iters, zero.n_iters)
  for out in (zero, learned):
    rel = abs(out.reg_ot_cost - gold["reg_ot_cost"]) / abs(gold["reg_ot_cost"])
    assert rel <= 1e-3, (out.reg_ot_cost, gold["reg_ot_cost"])


def test_gmm_d16_held_out_fewer_iterations():
  """4-component mixtures, n=m=2048, d=16, on seeds outside the training range."""
  from ott_amd.initializers.neural import train
  init = _initializer(16)
  iters_zero, iters_learned = [], []
  for seed in range(200000, 200016):
    xs, ys, _ = train.make_problem("gmm", seed, 2048, 2048, 16, eps_mult=0.05)
    x, y = torch.from_numpy(xs).cuda(), torch.from_numpy(ys).cuda()
    zero = _solve(x, y, None)
    learned = _solve(x, y, init)
    assert zero.converged and learned.converged
    # both runs converge on the same problem
    assert abs(learned.reg_ot_cost - zero.reg_ot_cost) <= 1e-3 * abs(zero.reg_ot_cost)
    iters_zero.append(zero.n_iters)
    iters_learned.append(learned.n_iters)
  assert np.mean(iters_learned) <= 0.9 * np.mean(iters_zero), (iters_learned, iters_zero)


def test_gmm_d16_learned_start_matches_fp64_oracle():
  """One held-out mixture problem: the warm-started BF16-split solve against the FP64 oracle."""
  from oracle import sinkhorn64
  from ott_amd.initializers.neural import train
  xs, ys, _ = train.make_problem("gmm", 200100, 2048, 2048, 16, eps_mult=0.05)
  out = _solve(torch.from_numpy(xs).cuda(), torch.from_numpy(ys).cuda(), _initializer(16))
  ref = sinkhorn64.solve_pointcloud(xs, ys, epsilon=0.05, relative_epsilon="mean", threshold=1e-3,
                                    inner_iterations=1, max_iterations=5000)
  assert out.converged and ref.converged
  rel = abs(out.reg_ot_cost - ref.reg_ot_cost) / abs(ref.reg_ot_cost)
  assert rel <= 1e-3, (out.reg_ot_cost, ref.reg_ot_cost)
