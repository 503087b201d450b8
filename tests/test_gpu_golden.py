"""Full-size GPU runs against the FP64 oracle fixtures in tests/golden/.

The fixtures hold the oracle's results (make_golden.py) for the benchmark
workloads: n=m=1024 d=2, n=m=16384 d=64, the unbalanced n=32768 m=8192 d=128
mixture (tau 0.95) and the batch of 512 uneven problems.  Checked: epsilon,
reg_ot_cost (rel 1e-4 fp32, 1e-3 bf16 split), iteration counts and the
recorded errors.
"""

import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"fp32": 1e-4, "bf16_split": 1e-3}


def _load(name):
  with open(os.path.join(HERE, name)) as fh:
    return json.load(fh)


def _solve(pb, precision, max_iterations, **kw):
  from ott_amd.geometry import pointcloud
  from ott_amd.solvers import linear
  geom = pointcloud.PointCloud(torch.from_numpy(pb.x).cuda(), torch.from_numpy(pb.y).cuda(),
                               epsilon=pb.eps_mult, relative_epsilon="mean", precision=precision)
  a = None if pb.a is None else torch.from_numpy(pb.a).cuda()
  b = None if pb.b is None else torch.from_numpy(pb.b).cuda()
  return linear.solve(geom, a=a, b=b, tau_a=pb.tau_a, tau_b=pb.tau_b, max_iterations=max_iterations, **kw), geom


def _check(out, geom, gold, precision):
  assert abs(geom.epsilon - gold["epsilon"]) <= 1e-6 * gold["epsilon"]
  rel = abs(out.reg_ot_cost - gold["reg_ot_cost"]) / abs(gold["reg_ot_cost"])
  assert rel <= TOL[precision], (out.reg_ot_cost, gold["reg_ot_cost"], rel)
  assert out.converged == gold["converged"]
  assert out.n_iters == gold["n_iters"], (out.n_iters, gold["n_iters"])
  errs = out.errors.cpu().numpy()[:len(gold["errors"])]
  np.testing.assert_allclose(errs, gold["errors"], rtol=20 * TOL[precision], atol=1e-6)
  f = out.f.cpu().numpy()
  g = out.g.cpu().numpy()
  # the gauge (f + c, g - c) is fixed by the iteration, so the means compare directly
  scale = max(1.0, abs(gold["f_mean"]))
  assert abs(f[np.isfinite(f)].mean() - gold["f_mean"]) <= 10 * TOL[precision] * scale
  assert abs(g[np.isfinite(g)].mean() - gold["g_mean"]) <= 10 * TOL[precision] * scale


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_uniform_shifted_n1024_d2(precision):
  from ott_amd import datasets
  gold = _load("uniform_shifted_n1024_d2.json")
  out, geom = _solve(datasets.uniform_shifted(1024, 1024, 2, seed=0), precision, 2000)
  _check(out, geom, gold, precision)
  with np.load(os.path.join(HERE, "uniform_shifted_n1024_d2.npz")) as z:
    f_ref, g_ref = z["f"], z["g"]
  tol = TOL[precision] * max(1.0, np.abs(f_ref).max())
  assert np.abs(out.f.cpu().numpy() - f_ref).max() <= tol
  assert np.abs(out.g.cpu().numpy() - g_ref).max() <= tol


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_gaussian_shifted_n16384_d64(precision):
  from ott_amd import datasets
  gold = _load("gaussian_shifted_n16384_d64.json")
  out, geom = _solve(datasets.gaussian_shifted(16384, 16384, 64, seed=1), precision, 5000)
  _check(out, geom, gold, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_mixture_pair_unbalanced_n32768_m8192_d128(precision):
  from ott_amd import datasets
  gold = _load("mixture_pair_n32768_m8192_d128.json")
  out, geom = _solve(datasets.mixture_pair(), precision, 2000)
  _check(out, geom, gold, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_batched_uneven_512_d32(precision):
  from ott_amd import datasets
  from ott_amd.tools import segment_sinkhorn as ss
  gold = _load("batched_uneven_512_d32.json")
  probs = datasets.batched_uneven()
  nx = [p.x.shape[0] for p in probs]
  ny = [p.y.shape[0] for p in probs]
  out = ss.segment_sinkhorn_outputs(
      torch.from_numpy(np.concatenate([p.x for p in probs])).cuda(),
      torch.from_numpy(np.concatenate([p.y for p in probs])).cuda(),
      num_per_segment_x=nx, num_per_segment_y=ny,
      sinkhorn_kwargs=dict(threshold=1e-3, inner_iterations=1, max_iterations=2000),
      epsilon=0.05, relative_epsilon="mean", padded_statistics=False, precision=precision)
  ref = np.array([g["reg_ot_cost"] for g in gold["problems"]])
  eps = np.array([g["epsilon"] for g in gold["problems"]])
  iters = np.array([g["n_iters"] for g in gold["problems"]])
  np.testing.assert_allclose(out.epsilon.numpy(), eps, rtol=1e-6)
  rel = np.abs(out.reg_ot_cost.numpy() - ref) / np.abs(ref)
  assert rel.max() <= TOL[precision], rel.max()
  assert bool(out.converged.all())
  # a problem whose error lands within float32 noise of the threshold may stop one iteration apart
  diff = np.abs(out.n_iters.numpy() - iters)
  assert diff.max() <= 1 and (diff > 0).sum() <= 5, (diff.max(), (diff > 0).sum())
