"""Output-side transport operations of SinkhornOutput (sinkhorn.py:337-470)
against the dense FP64 oracle, evaluated at the potentials the GPU solve
returned: ``matrix`` (geometry.py:361-365), ``apply`` both ways with signed and
batched inputs (geometry.py:445-486), ``primal_cost`` and
``transport_cost_at_geom`` (sinkhorn.py:345-411; one fused pass over the pairs,
here with another cloud pair of 37 features, more than one staging chunk)."""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu


def _solved(precision):
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  rng = np.random.default_rng(31)
  x = (rng.normal(size=(300, 8)) / np.sqrt(8)).astype(np.float32)
  y = (rng.normal(size=(200, 8)) / np.sqrt(8) + 0.25).astype(np.float32)
  a = rng.uniform(0.2, 1.2, 300)
  a[7] = 0.0
  a = (a / a.sum()).astype(np.float32)
  b = rng.uniform(0.2, 1.2, 200)
  b = (b / b.sum()).astype(np.float32)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=0.05,
                               relative_epsilon="mean", precision=precision)
  prob = linear_problem.LinearProblem(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
  out = sinkhorn.Sinkhorn(threshold=1e-5, inner_iterations=1)(prob)
  assert out.converged
  f = out.f.double().cpu().numpy()
  g = out.g.double().cpu().numpy()
  cost = oracle.sqeuclidean_cost(x.astype(np.float64), y.astype(np.float64))
  return out, x, y, f, g, cost, geom.epsilon


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_matrix(precision):
  out, x, y, f, g, cost, eps = _solved(precision)
  p_ref = oracle.transport_from_potentials(cost, f, g, eps)
  p = out.matrix.double().cpu().numpy()
  assert p.shape == (300, 200)
  assert np.all(p[7] == 0.0)  # zero weight: f = -inf
  np.testing.assert_allclose(p, p_ref, rtol=1e-4, atol=1e-6 * p_ref.max())


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("axis", [0, 1])
def test_apply_signed_batched(precision, axis):
  out, x, y, f, g, cost, eps = _solved(precision)
  rng = np.random.default_rng(2)
  vec = rng.normal(size=(3, 300 if axis == 0 else 200)).astype(np.float32)
  vec[1] = np.abs(vec[1])  # one vector of a single sign
  got = out.apply(torch.from_numpy(vec).cuda(), axis=axis).double().cpu().numpy()
  ref = oracle.apply_transport_from_potentials(cost, f, g, vec.astype(np.float64), eps, axis=axis)
  # the scale of the sum of absolute terms bounds the cancellation
  mag = oracle.apply_transport_from_potentials(cost, f, g, np.abs(vec).astype(np.float64), eps, axis=axis)
  tol = 2e-5 if precision == "fp32" else 1e-4
  np.testing.assert_allclose(got, ref, atol=tol * float(mag.max()))
  one = out.apply(torch.from_numpy(vec[0]).cuda(), axis=axis).double().cpu().numpy()
  np.testing.assert_allclose(one, got[0], rtol=0, atol=1e-12 * float(mag.max()) + 1e-30)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("d", [8, 100])
def test_apply_many_vectors(precision, axis, d):
  """37 signed vectors (launches of 16, 16 and 5) on a problem with several
  row splits, wide features, a zero weight, an all-zero vector and a NaN
  entry, which makes its vector's output NaN as the reference's signed
  log-sum-exp does."""
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  rng = np.random.default_rng(40 + d)
  n, m = 1500, 1100
  x = (rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32)
  y = (rng.normal(size=(m, d)) / np.sqrt(d) + 0.2 / np.sqrt(d)).astype(np.float32)
  a = rng.uniform(0.2, 1.2, n)
  a[11] = 0.0
  a = (a / a.sum()).astype(np.float32)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=0.05,
                               relative_epsilon="mean", precision=precision)
  out = sinkhorn.Sinkhorn(threshold=1e-4)(linear_problem.LinearProblem(geom, torch.from_numpy(a).cuda()))
  assert out.converged
  f, g = out.f.double().cpu().numpy(), out.g.double().cpu().numpy()
  cost = oracle.sqeuclidean_cost(x.astype(np.float64), y.astype(np.float64))
  vec = rng.normal(size=(37, n if axis == 0 else m)).astype(np.float32)
  vec[9] = 0.0
  vec[20, 3] = np.nan
  got = out.apply(torch.from_numpy(vec).cuda(), axis=axis).double().cpu().numpy()
  ok = np.ones(37, bool)
  ok[20] = False
  ref = oracle.apply_transport_from_potentials(cost, f, g, vec[ok].astype(np.float64), geom.epsilon, axis=axis)
  mag = oracle.apply_transport_from_potentials(cost, f, g, np.abs(vec[ok]).astype(np.float64), geom.epsilon,
                                               axis=axis)
  # bf16 split: the cross term's relative error times the exponents' size
  tol = 2e-5 if precision == "fp32" else 1e-3
  np.testing.assert_allclose(got[ok], ref, atol=tol * float(mag.max()))
  assert np.all(got[9] == 0.0)
  assert np.all(np.isnan(got[20]))


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_primal_cost_and_transport_cost_at_geom(precision):
  from ott_amd.geometry import pointcloud
  out, x, y, f, g, cost, eps = _solved(precision)
  ref = oracle.transport_cost_at_cost(cost, f, g, eps, cost)
  tol = 1e-5 if precision == "fp32" else 1e-4
  assert math.isclose(out.primal_cost, ref, rel_tol=tol), (out.primal_cost, ref)
  rng = np.random.default_rng(4)
  x2 = (rng.normal(size=(300, 37)) + 1.0).astype(np.float32)
  y2 = rng.normal(size=(200, 37)).astype(np.float32)
  other = pointcloud.PointCloud(torch.from_numpy(x2).cuda(), torch.from_numpy(y2).cuda(), epsilon=1.0,
                                scale_cost=2.0)
  other_cost = oracle.sqeuclidean_cost(x2.astype(np.float64), y2.astype(np.float64), 0.5)
  ref2 = oracle.transport_cost_at_cost(cost, f, g, eps, other_cost)
  assert math.isclose(out.transport_cost_at_geom(other), ref2, rel_tol=tol), (out.transport_cost_at_geom(other), ref2)
  # at convergence the primal cost and eps KL(P | a b^T) add up to reg_ot_cost (sinkhorn.py:224-277, :350-384)
  p = oracle.transport_from_potentials(cost, f, g, eps)
  a = out.a.double().cpu().numpy()
  b = out.b.double().cpu().numpy()
  q = a[:, None] * b[None, :]
  with np.errstate(divide="ignore", invalid="ignore"):
    kl = float(np.where(p > 0, p * np.log(p / q), 0.0).sum() - p.sum() + q.sum())
  assert math.isclose(out.primal_cost + eps * kl, out.reg_ot_cost, rel_tol=1e-3)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_entropy_and_cost_at_geom(precision):
  """entropy (sinkhorn.py:512-522) is sum(entr(P)) and cost_at_geom
  (sinkhorn.py:474-479) is <P, C'> - eps sum(entr(P)), both against the dense
  matrix of the oracle."""
  from ott_amd.geometry import pointcloud
  from scipy import special
  out, x, y, f, g, cost, eps = _solved(precision)
  p = oracle.transport_from_potentials(cost, f, g, eps)
  ent = float(special.entr(p).sum())
  tol = 1e-5 if precision == "fp32" else 1e-4
  assert math.isclose(out.entropy, ent, rel_tol=tol), (out.entropy, ent)
  rng = np.random.default_rng(8)
  x2 = rng.normal(size=(300, 3)).astype(np.float32)
  y2 = rng.normal(size=(200, 3)).astype(np.float32)
  other = pointcloud.PointCloud(torch.from_numpy(x2).cuda(), torch.from_numpy(y2).cuda(), epsilon=1.0)
  other_cost = oracle.sqeuclidean_cost(x2.astype(np.float64), y2.astype(np.float64))
  ref = float((p * other_cost).sum()) - eps * ent
  got = out.cost_at_geom(other)
  assert abs(got - ref) <= tol * float((p * other_cost).sum()), (got, ref)


def test_diag_and_normalized_entropy():
  """diag (sinkhorn.py:531-538) and normalized_entropy (:524-529) of an
  assignment problem (n = m, uniform weights)."""
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  from scipy import special
  rng = np.random.default_rng(17)
  x = rng.normal(size=(128, 4)).astype(np.float32)
  y = (rng.normal(size=(128, 4)) + 0.1).astype(np.float32)
  geom = pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), epsilon=0.5)
  out = sinkhorn.Sinkhorn(threshold=1e-5)(linear_problem.LinearProblem(geom))
  assert out.converged
  f = out.f.double().cpu().numpy()
  g = out.g.double().cpu().numpy()
  p = oracle.transport_from_potentials(oracle.sqeuclidean_cost(x.astype(np.float64), y.astype(np.float64)), f, g, 0.5)
  np.testing.assert_allclose(out.diag.double().cpu().numpy(), np.diag(p), rtol=1e-5, atol=1e-12)
  ent = float(special.entr(p).sum())
  assert math.isclose(out.normalized_entropy, ent / math.log(128) - 1.0, rel_tol=1e-5, abs_tol=1e-6)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_to_dual_potentials(precision):
  """SinkhornOutput.to_dual_potentials (sinkhorn.py:480-499): f and g as soft
  c-transforms at new points (batched and single, g with the reference's
  eps log k for a batch of k), the forward and backward transport maps
  (potentials.py:52-84) and the distance (potentials.py:86-101), against the
  oracle."""
  out, x, y, f, g, cost, eps = _solved(precision)
  a = out.a.double().cpu().numpy()
  b = out.b.double().cpu().numpy()
  dp = out.to_dual_potentials()
  rng = np.random.default_rng(12)
  xs = (rng.normal(size=(50, 8)) / np.sqrt(8)).astype(np.float32)
  ys = (rng.normal(size=(40, 8)) / np.sqrt(8) + 0.25).astype(np.float32)
  cx = oracle.sqeuclidean_cost(xs.astype(np.float64), y.astype(np.float64))
  cy = oracle.sqeuclidean_cost(x.astype(np.float64), ys.astype(np.float64))
  f_ref = oracle.soft_c_transform(cx, g, b, eps, axis=1)
  g_ref = oracle.soft_c_transform(cy, f, np.full(40, 1.0 / 40), eps, axis=0)
  tol = 1e-4 if precision == "fp32" else 1e-3
  scale = max(np.abs(f_ref).max(), np.abs(g_ref).max(), eps)
  np.testing.assert_allclose(dp.f(torch.from_numpy(xs).cuda()).double().cpu().numpy(), f_ref, atol=tol * scale)
  np.testing.assert_allclose(dp.g(torch.from_numpy(ys).cuda()).double().cpu().numpy(), g_ref, atol=tol * scale)
  g_one = oracle.soft_c_transform(cy[:, :1], f, np.ones(1), eps, axis=0)[0]
  assert abs(float(dp.g(torch.from_numpy(ys[0]).cuda())) - g_one) <= tol * scale
  t_fwd = oracle.barycentric_map(cx, g, b, eps, y)
  t_bwd = oracle.barycentric_map(cy.T, f, np.ones(len(f)), eps, x)
  np.testing.assert_allclose(dp.transport(torch.from_numpy(xs).cuda()).double().cpu().numpy(), t_fwd,
                             atol=tol * np.abs(y).max())
  np.testing.assert_allclose(dp.transport(torch.from_numpy(ys).cuda(), forward=False).double().cpu().numpy(), t_bwd,
                             atol=tol * np.abs(x).max())
  g_each = np.array([oracle.soft_c_transform(cy[:, j:j + 1], f, np.ones(1), eps, axis=0)[0] for j in range(40)])
  dist_ref = f_ref.mean() + g_each.mean()
  assert abs(dp.distance(torch.from_numpy(xs).cuda(), torch.from_numpy(ys).cuda()) - dist_ref) <= tol * scale
