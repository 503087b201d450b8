"""Costs other than the squared Euclidean (geometry/costs.py PNormP, SqPNorm,
EuclideanP, Euclidean) on the GPU against the FP64 oracle's dense cost
matrices (``oracle.ti_cost``).

Those with p != 2 run the vector-ALU cost stream inside the fused half-step
(summed from the differences, float32); PNormP(2), SqPNorm(2) and
EuclideanP(2) are (half) the squared Euclidean cost and take its MFMA path.
Tolerances as the parity tests: 1e-4 for fp32, 1e-3 for the bf16 split (which
only the MFMA-routed costs use).
"""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle
from tests.test_gpu_parity import TOL, _check_solve, _close, _cloud

pytestmark = pytest.mark.gpu

SPECS = [("pnormp", 1.0), ("pnormp", 1.5), ("sqpnorm", 1.0), ("sqpnorm", 3.0), ("euclideanp", 1.0),
         ("euclideanp", 3.0), ("euclidean",), ("pnormp", 2.0), ("sqpnorm", 2.0), ("euclideanp", 2.0)]


def _cost(spec):
  from ott_amd.geometry import costs
  kind = spec[0]
  if kind == "euclidean":
    return costs.Euclidean()
  return {"pnormp": costs.PNormP, "sqpnorm": costs.SqPNorm, "euclideanp": costs.EuclideanP}[kind](spec[1])


def _geom(x, y, spec, precision="fp32", **kw):
  from ott_amd.geometry import pointcloud
  return pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), cost_fn=_cost(spec),
                               precision=precision, **kw)


def _tol(spec, precision):
  # only the squared-Euclidean kinds run the bf16 split
  mfma = spec[0] != "euclidean" and spec[1] == 2.0
  return TOL[precision if mfma else "fp32"]


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: "-".join(map(str, s)))
@pytest.mark.parametrize("axis", [0, 1])
def test_apply_lse_kernel_other_costs(spec, axis):
  """apply_lse_kernel (geometry.py:260-299) on the cost stream, rows and
  columns past a 64-point tile, d above the 32-feature staging chunk, a -inf
  potential."""
  x, y, _, _ = _cloud(331, 203, 37, seed=3)
  rng = np.random.default_rng(5)
  f = (rng.normal(size=331) * 0.1).astype(np.float32)
  g = (rng.normal(size=203) * 0.1).astype(np.float32)
  f[7] = -np.inf
  eps = 0.3
  print(f"apply_lse {spec} axis={axis}", flush=True)
  geom = _geom(x, y, spec, epsilon=eps)
  got, _ = geom.apply_lse_kernel(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(), eps, axis=axis)
  ops = oracle.DenseLSE(oracle.ti_cost(x, y, spec))
  want = oracle.apply_lse_kernel(ops, f.astype(np.float64), g.astype(np.float64), eps, axis)
  _close(got.cpu().numpy(), want, 1e-4, f"apply_lse {spec}")


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: "-".join(map(str, s)))
@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_solve_other_costs(spec, precision):
  """A full solve (linear.solve) against the oracle's solve on the dense cost,
  absolute epsilon, zero weights on both sides."""
  from ott_amd.solvers import linear
  x, y, a, b = _cloud(300, 257, 6, seed=12, zero_a=[4], zero_b=[9])
  print(f"solve {spec} {precision}", flush=True)
  eps = 0.05 * float(np.mean(oracle.ti_cost(x, y, spec)))
  geom = _geom(x, y, spec, precision, epsilon=eps)
  kw = dict(threshold=1e-4, inner_iterations=5, max_iterations=3000)
  out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, cost_fn=spec, **kw)
  assert ref.converged
  tol = _tol(spec, precision)
  _check_solve(out, ref, "fp32" if tol == TOL["fp32"] else "bf16_split")


@pytest.mark.parametrize("spec", [("pnormp", 1.0), ("euclidean",), ("sqpnorm", 3.0)],
                         ids=lambda s: "-".join(map(str, s)))
@pytest.mark.parametrize("mode", ["mean", "std", "default"])
def test_relative_epsilon_other_costs(spec, mode):
  """relative_epsilon "mean" / "std" and the default 0.05 std (geometry.py:
  147-171) from the GPU's float64 sums over the pairs, then the solve."""
  from ott_amd.solvers import linear
  x, y, a, b = _cloud(211, 190, 5, seed=21)
  c = oracle.ti_cost(x, y, spec)
  if mode == "default":
    geom = _geom(x, y, spec)
    want = 0.05 * float(np.std(c))
    ref_kw = dict(epsilon=None, relative_epsilon=None)
  else:
    geom = _geom(x, y, spec, epsilon=0.1, relative_epsilon=mode)
    want = 0.1 * float(np.mean(c) if mode == "mean" else np.std(c))
    ref_kw = dict(epsilon=0.1, relative_epsilon=mode)
  assert math.isclose(geom.epsilon, want, rel_tol=1e-6), (geom.epsilon, want)
  assert math.isclose(geom.mean_cost_matrix, float(np.mean(c)), rel_tol=1e-6)
  assert math.isclose(geom.std_cost_matrix, float(np.std(c)), rel_tol=1e-5)
  kw = dict(threshold=1e-4, inner_iterations=1, max_iterations=3000)
  out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, cost_fn=spec, **ref_kw, **kw)
  _check_solve(out, ref, "fp32")


@pytest.mark.parametrize("spec", [("pnormp", 1.0), ("euclideanp", 1.5), ("sqpnorm", 2.0)],
                         ids=lambda s: "-".join(map(str, s)))
@pytest.mark.parametrize("mode", ["mean", "max_cost", "median"])
def test_scale_cost_modes_other_costs(spec, mode):
  """scale_cost "mean", "max_cost", "median" (pointcloud.py:342-383) of the
  unscaled cost matrix, then a solve on the scaled cost."""
  from ott_amd.solvers import linear
  x, y, a, b = _cloud(233, 171, 4, seed=29, shift=0.4)
  c = oracle.ti_cost(x, y, spec)
  want = {"mean": np.mean(c), "max_cost": np.max(c), "median": np.median(c)}[mode]
  geom = _geom(x, y, spec, epsilon=0.05, relative_epsilon="mean", scale_cost=mode)
  assert math.isclose(1.0 / geom.inv_scale_cost, float(want), rel_tol=2e-6), (1.0 / geom.inv_scale_cost, want)
  kw = dict(threshold=1e-4, inner_iterations=1, max_iterations=3000)
  out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, relative_epsilon="mean",
                                scale_cost=1.0 / geom.inv_scale_cost, cost_fn=spec, **kw)
  _check_solve(out, ref, "fp32")


@pytest.mark.parametrize("spec", [("pnormp", 1.0), ("euclidean",), ("euclideanp", 3.0)],
                         ids=lambda s: "-".join(map(str, s)))
def test_unbalanced_options_other_costs(spec):
  """Unbalanced with momentum and the recentred iteration on the cost stream."""
  from ott_amd.solvers import linear
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(190, 143, 5, seed=31, zero_b=[2])
  eps = 0.05 * float(np.mean(oracle.ti_cost(x, y, spec)))
  kw = dict(threshold=1e-4, inner_iterations=3, max_iterations=3000, tau_a=0.9, tau_b=0.8)
  for rc in (False, True):
    geom = _geom(x, y, spec, epsilon=eps)
    out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), recenter_potentials=rc,
                       momentum=acceleration.Momentum(value=1.2), **kw)
    ref = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, cost_fn=spec, recenter_potentials=rc, momentum=1.2,
                                  **kw)
    assert ref.converged
    _check_solve(out, ref, "fp32")


@pytest.mark.parametrize("spec", [("pnormp", 1.0), ("sqpnorm", 3.0), ("euclidean",)],
                         ids=lambda s: "-".join(map(str, s)))
def test_transport_outputs_other_costs(spec):
  """matrix, primal_cost and the median of the cost matrix on the cost
  stream's costs; the operations built on the squared-Euclidean cross term
  refuse."""
  from ott_amd.solvers import linear
  x, y, a, b = _cloud(150, 97, 5, seed=37)
  c = oracle.ti_cost(x, y, spec)
  eps = 0.05 * float(np.mean(c))
  geom = _geom(x, y, spec, epsilon=eps)
  out = linear.solve(geom, a=torch.from_numpy(a).cuda(), b=torch.from_numpy(b).cuda(), threshold=1e-5)
  f = out.f.double().cpu().numpy()
  g = out.g.double().cpu().numpy()
  p_ref = oracle.transport_from_potentials(c, f, g, eps)
  np.testing.assert_allclose(out.matrix.double().cpu().numpy(), p_ref, rtol=2e-4, atol=1e-9)
  want = oracle.transport_cost_at_cost(c, f, g, eps, c)
  assert math.isclose(out.primal_cost, want, rel_tol=1e-4), (out.primal_cost, want)
  assert math.isclose(geom.median_cost_matrix, float(np.median(c)), rel_tol=2e-6)
  with pytest.raises(NotImplementedError):
    out.apply(torch.ones(150, device="cuda"))
  with pytest.raises(NotImplementedError):
    out.reg_ot_cost_grad()
