"""Pins the FP64 oracle with checks that do not reuse its own formulas.

The reference (JAX) cannot run here and its tests hold no golden values for
this path, so each oracle function is checked against something independent:
brute force, scipy, closed forms, an L-BFGS solve of the semi-dual, the
Hungarian and epsilon -> infinity limits, and invariances.  A plausible slip
(dropped norm term, wrong sign, swapped axis, wrong error cadence) fails one of
these.
"""

import math

import numpy as np
import pytest
from scipy import optimize, special

from oracle import sinkhorn64 as oracle


def _problem(n, m, d, seed, zero_a=None, zero_b=None, shift=0.25):
  rng = np.random.default_rng(seed)
  x = rng.uniform(size=(n, d))
  y = rng.uniform(size=(m, d)) + shift
  a = rng.uniform(size=n) + 0.2
  b = rng.uniform(size=m) + 0.2
  if zero_a is not None:
    a[zero_a] = 0.0
  if zero_b is not None:
    b[zero_b] = 0.0
  return x, y, a / a.sum(), b / b.sum()


def test_cost_brute_force():
  x, y, _, _ = _problem(7, 5, 3, 0)
  want = np.array([[sum((x[i, k] - y[j, k]) ** 2 for k in range(3)) for j in range(5)]
                   for i in range(7)])
  np.testing.assert_allclose(oracle.sqeuclidean_cost(x, y), want, rtol=1e-14)
  np.testing.assert_allclose(oracle.sqeuclidean_cost(x, y, expanded=True), want, rtol=1e-12)
  np.testing.assert_allclose(oracle.sqeuclidean_cost(x, y, 0.5), 0.5 * want, rtol=1e-14)


def test_mean_cost_and_epsilon():
  x, y, _, _ = _problem(40, 30, 4, 1)
  c = oracle.sqeuclidean_cost(x, y)
  assert math.isclose(oracle.mean_cost(x, y), c.mean(), rel_tol=1e-12)
  assert math.isclose(oracle.std_cost(x, y), c.std(), rel_tol=1e-10)
  assert math.isclose(oracle.resolve_epsilon(x, y, 0.05, "mean"), 0.05 * c.mean(), rel_tol=1e-12)
  assert math.isclose(oracle.resolve_epsilon(x, y, None, None), 0.05 * c.std(), rel_tol=1e-10)
  assert oracle.resolve_epsilon(x, y, 0.3, None) == 0.3
  assert oracle.epsilon_at(2.0, 3, init=10.0, decay=0.5) == max(10 * 0.125, 1.0) * 2.0


def test_lse_naive_and_neginf():
  rng = np.random.default_rng(2)
  z = rng.normal(size=(6, 9))
  np.testing.assert_allclose(oracle._lse(z, 0), np.log(np.exp(z).sum(0)), rtol=1e-13)
  np.testing.assert_allclose(oracle._lse(z, 1), np.log(np.exp(z).sum(1)), rtol=1e-13)
  z[:, 2] = -np.inf
  assert np.isneginf(oracle._lse(z, 0)[2])


def test_half_step_closed_form_m1():
  """m = 1: g = eps log b - eps LSE_i((f_i - C_i) / eps), written with math.exp."""
  x, y, a, _ = _problem(9, 1, 2, 3)
  b = np.array([1.0])
  f = np.linspace(-0.3, 0.2, 9)
  eps = 0.07
  c = [sum((x[i, k] - y[0, k]) ** 2 for k in range(2)) for i in range(9)]
  s = sum(math.exp((f[i] - c[i]) / eps) for i in range(9))
  want = eps * math.log(1.0) - eps * math.log(s)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  got = oracle.update_potential(ops, f, np.zeros(1), np.log(b), eps, axis=0)
  assert math.isclose(got[0], want, rel_tol=1e-12, abs_tol=1e-14)


@pytest.mark.parametrize("axis", [0, 1])
def test_half_step_gives_exact_marginal(axis):
  x, y, a, b = _problem(13, 11, 3, 4, zero_a=[2], zero_b=[5])
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  eps = 0.05
  rng = np.random.default_rng(0)
  f = np.where(a > 0, rng.normal(size=13) * 0.1, -np.inf)
  g = np.where(b > 0, rng.normal(size=11) * 0.1, -np.inf)
  with np.errstate(divide="ignore"):
    if axis == 0:
      g = oracle.update_potential(ops, f, g, np.log(b), eps, 0)
      target = b
    else:
      f = oracle.update_potential(ops, f, g, np.log(a), eps, 1)
      target = a
  marg = oracle.marginal_from_potentials(ops, f, g, eps, axis)
  np.testing.assert_allclose(marg, target, rtol=1e-12, atol=1e-15)
  # and the marginal equals the row/column sums of the explicit plan
  plan = np.exp((f[:, None] + g[None, :] - ops.cost) / eps)
  np.testing.assert_allclose(marg, plan.sum(axis=axis), rtol=1e-12, atol=1e-15)


def test_fixed_point_matches_lbfgs_semidual():
  x, y, a, b = _problem(12, 9, 2, 5)
  eps = 0.1
  c = oracle.sqeuclidean_cost(x, y)
  out = oracle.sinkhorn(oracle.DenseLSE(c), a, b, eps, threshold=1e-12, max_iterations=20000)
  assert out.converged

  def neg_semidual(gv):
    f = eps * np.log(a) - eps * special.logsumexp((gv[None, :] - c) / eps, axis=1)
    val = a @ f + b @ gv
    p = np.exp((f[:, None] + gv[None, :] - c) / eps)
    grad = b - p.sum(0)
    return -val, -grad

  res = optimize.minimize(neg_semidual, np.zeros(9), jac=True, method="L-BFGS-B",
                          options=dict(gtol=1e-13, ftol=1e-15, maxiter=5000))
  g_star = res.x - res.x.mean()
  g_or = out.g - out.g.mean()
  np.testing.assert_allclose(g_or, g_star, atol=1e-6)
  # At the optimum the semi-dual value plus eps (H(a) + H(b)) is reg_ot_cost:
  # OTT's potentials carry eps log a, eps log b (KL relative to a b^T).
  ent = -(a @ np.log(a)) - (b @ np.log(b))
  assert math.isclose(out.reg_ot_cost, -res.fun + eps * ent, rel_tol=1e-8)


def test_hungarian_limit():
  rng = np.random.default_rng(6)
  n = 8
  x = rng.uniform(size=(n, 2))
  y = rng.uniform(size=(n, 2))
  c = oracle.sqeuclidean_cost(x, y)
  a = b = np.full(n, 1.0 / n)
  out = oracle.sinkhorn(oracle.DenseLSE(c), a, b, 0.002, threshold=1e-6, max_iterations=50000)
  plan = np.exp((out.f[:, None] + out.g[None, :] - c) / 0.002)
  rows, cols = optimize.linear_sum_assignment(c)
  hung = np.zeros_like(plan)
  hung[rows, cols] = 1.0 / n
  # sinkhorn_test.py:68-90 criterion: cosine distance to the Hungarian plan < 0.2
  cos = plan.ravel() @ hung.ravel() / (np.linalg.norm(plan) * np.linalg.norm(hung))
  assert 1.0 - cos < 0.2
  assert plan[rows, cols].sum() > 0.95
  # sinkhorn_test.py:56-66: x against itself at small eps puts > 0.99 on the
  # diagonal (17 points in [0, 1]^4 as in the reference; eps 0.002 because
  # these numpy-seeded points are not the reference's jax-seeded ones)
  x = rng.uniform(size=(17, 4))
  a = np.full(17, 1.0 / 17)
  cxx = oracle.sqeuclidean_cost(x, x)
  out = oracle.sinkhorn(oracle.DenseLSE(cxx), a, a, 0.002, threshold=1e-6, max_iterations=50000)
  diag = np.exp((out.f + out.g - np.diag(cxx)) / 0.002)
  assert 0.99 < diag.sum() < 1.0


def test_large_epsilon_limit():
  x, y, a, b = _problem(10, 7, 3, 7)
  c = oracle.sqeuclidean_cost(x, y)
  out = oracle.sinkhorn(oracle.DenseLSE(c), a, b, 1e4, threshold=1e-12)
  assert math.isclose(out.reg_ot_cost, float(a @ c @ b), rel_tol=1e-5)


def test_scale_invariance():
  x, y, a, b = _problem(15, 12, 3, 8)
  s = 3.0
  o1 = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, threshold=1e-10, max_iterations=10000)
  o2 = oracle.solve_pointcloud(s * x, s * y, a, b, epsilon=0.05 * s * s, threshold=1e-10,
                               max_iterations=10000)
  np.testing.assert_allclose(o2.f, s * s * o1.f, rtol=1e-6, atol=1e-9)
  assert math.isclose(o2.reg_ot_cost, s * s * o1.reg_ot_cost, rel_tol=1e-8)


def test_gauge_invariance_of_cost():
  x, y, a, b = _problem(11, 10, 2, 9, zero_a=[1], zero_b=[4])
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  out = oracle.sinkhorn(ops, a, b, 0.1, threshold=1e-9)
  r1 = oracle.compute_kl_reg_cost(ops, out.f, out.g, a, b, 0.1)
  r2 = oracle.compute_kl_reg_cost(ops, out.f + 0.3, out.g - 0.3, a, b, 0.1)
  assert math.isclose(r1, r2, rel_tol=1e-12)
  assert np.isneginf(out.f[1]) and np.isneginf(out.g[4])
  assert np.isfinite(out.reg_ot_cost)


def test_unbalanced_scalar_fixed_point():
  """n = m = 1: f = ta (A - g), g = tb (B - f) has a closed-form fixed point."""
  x = np.array([[0.0, 0.0]])
  y = np.array([[0.6, 0.8]])
  a = np.array([0.7])
  b = np.array([1.3])
  eps, ta, tb = 0.2, 0.8, 0.9
  c0 = 1.0
  A = eps * math.log(0.7) + c0
  B = eps * math.log(1.3) + c0
  f_star = ta * (A - tb * B) / (1 - ta * tb)
  g_star = tb * (B - f_star)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  out = oracle.sinkhorn(ops, a, b, eps, tau_a=ta, tau_b=tb, threshold=1e-13, max_iterations=5000)
  assert math.isclose(out.f[0], f_star, rel_tol=1e-9)
  assert math.isclose(out.g[0], g_star, rel_tol=1e-9)


def test_loop_bookkeeping():
  x, y, a, b = _problem(17, 29, 4, 10)
  kw = dict(epsilon=0.1)
  out = oracle.solve_pointcloud(x, y, a, b, min_iterations=34, **kw)
  assert np.isinf(out.errors[:3]).all() and out.errors[3] > 0
  assert out.converged
  assert out.n_iters == int(np.sum(out.errors != -1)) * 10
  out = oracle.solve_pointcloud(x, y, a, b, min_iterations=50, max_iterations=50, **kw)
  assert out.converged and 0 < out.errors[-1] < 1e-3 and np.isinf(out.errors[:-1]).all()
  out = oracle.solve_pointcloud(x, y, a, b, max_iterations=25, threshold=1e-12, **kw)
  assert out.errors.shape == (3,) and not out.converged and out.n_iters == 30
  # a min_iterations multiple of inner stops on the +inf of the previous block
  out = oracle.solve_pointcloud(x, y, a, b, min_iterations=40, **kw)
  assert out.n_iters == 40 and np.isinf(out.errors[3]) and out.errors[4] == -1
  # progress callback once per iteration
  calls = []
  oracle.solve_pointcloud(x, y, a, b, progress_fn=lambda s: calls.append(s[0]), **kw)
  assert calls == list(range(len(calls))) and len(calls) % 10 == 0


def test_unbalanced_converges_and_semibalanced():
  x, y, a, b = _problem(30, 20, 3, 11)
  for ta, tb in [(0.8, 0.9), (1.0, 0.9)]:
    out = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, tau_a=ta, tau_b=tb)
    err = out.errors[out.errors > -1][-1]
    assert out.converged and 0 < err < 1e-3


def test_c_oracle_matches_dense():
  x, y, a, b = _problem(130, 70, 5, 12, zero_b=[3])
  dense = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  online = oracle.PointCloudLSE(x, y)
  rng = np.random.default_rng(1)
  f = rng.normal(size=130) * 0.1
  g = rng.normal(size=70) * 0.1
  g[3] = -np.inf
  for axis in (0, 1):
    want = dense.lse(f, g, 0.05, axis)
    got = online.lse(f, g, 0.05, axis)
    fin = np.isfinite(want)
    np.testing.assert_allclose(got[fin], want[fin], rtol=1e-12)
    assert np.array_equal(np.isfinite(got), fin)


def test_fp32_twin_agrees():
  x, y, a, b = _problem(64, 48, 3, 13)
  o64 = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05)
  o32 = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, dtype=np.float32, expanded=True)
  assert o32.n_iters == o64.n_iters
  np.testing.assert_allclose(o32.f, o64.f, atol=2e-5)
  assert math.isclose(o32.reg_ot_cost, o64.reg_ot_cost, rel_tol=1e-5)


def test_oracle_recenter_balanced_closed_form():
  """SinkhornState.recenter, balanced (sinkhorn.py:98-102): zero mean over the
  finite f, f_i + g_j unchanged, -inf entries kept."""
  from oracle import sinkhorn64 as o
  rng = np.random.default_rng(0)
  f = rng.normal(size=9)
  g = rng.normal(size=5)
  f[3] = -np.inf
  a = np.full(9, 1 / 9.0)
  b = np.full(5, 0.2)
  fr, gr = o.recenter(f, g, a, b, 0.1, 1.0, 1.0)
  fin = np.isfinite(f)
  assert abs(fr[fin].mean()) < 1e-14 and np.isneginf(fr[3])
  np.testing.assert_allclose((fr[:, None] + gr[None])[fin], (f[:, None] + g[None])[fin], atol=1e-14)
  np.testing.assert_allclose(fr[fin], f[fin] - f[fin].mean(), atol=1e-14)
  # one tau = 1: unchanged (sinkhorn.py:104-106)
  f2, g2 = o.recenter(f, g, a, b, 0.1, 1.0, 0.9)
  assert f2 is f and g2 is g


def test_oracle_parallel_updates_on_a_symmetric_problem():
  """Parallel updates with momentum 0.5 on x against x with a = b (the
  symmetric terms of sinkhorn_divergence.py:298-318): f = g at every
  iteration (equal starts, a symmetric cost), the same fixed point as the
  sequential solver, and the parallel error adds the row and the column
  marginal errors (sinkhorn.py:160-189), which are equal here."""
  rng = np.random.default_rng(11)
  x = rng.normal(size=(40, 3))
  a = rng.uniform(0.2, 1.0, size=40)
  a /= a.sum()
  kw = dict(epsilon=0.1, threshold=1e-9, inner_iterations=1, max_iterations=5000)
  par = oracle.solve_pointcloud(x, x, a, a, parallel_dual_updates=True, momentum=0.5, **kw)
  # sequential updates oscillate on a symmetric problem and crawl: 5000
  # iterations leave them near, not at, the fixed point the parallel ones
  # reach in a few dozen
  seq = oracle.solve_pointcloud(x, x, a, a, **kw)
  assert par.converged and par.n_iters < 100
  np.testing.assert_allclose(par.f, par.g, atol=1e-12)  # equal up to rounding
  assert math.isclose(par.reg_ot_cost, seq.reg_ot_cost, rel_tol=1e-7)
  # the symmetric fixed point: f = eps log a - eps LSE_j((f_j - C_ij) / eps)
  cost = ((x[:, None, :] - x[None]) ** 2).sum(-1)
  t = 0.1 * np.log(a) - 0.1 * special.logsumexp((par.f[None, :] - cost) / 0.1, axis=1)
  np.testing.assert_allclose(par.f, t, atol=1e-8)
  # the error of the first iteration: twice the L1 distance of the column marginal to a
  one = oracle.solve_pointcloud(x, x, a, a, parallel_dual_updates=True, momentum=0.5, epsilon=0.1,
                                threshold=0.0, inner_iterations=1, max_iterations=1)
  plan = np.exp((one.f[:, None] + one.g[None] - cost) / 0.1)
  np.testing.assert_allclose(one.errors[0], 2.0 * np.abs(plan.sum(0) - a).sum(), rtol=1e-10)


def test_oracle_transport_ops_against_reg_ot_cost():
  """At convergence the dual value equals the primal one:
  reg_ot_cost = <P, C> + eps KL(P | a b^T) with KL(P|Q) = sum P log(P/Q) - P + Q
  (sinkhorn.py:224-277 against :350-384), which pins the transport matrix and
  the transport cost of the oracle; apply is the matrix product both ways."""
  rng = np.random.default_rng(5)
  x = rng.normal(size=(30, 3))
  y = rng.normal(size=(25, 3)) + 0.4
  a = rng.uniform(0.3, 1.0, 30)
  a /= a.sum()
  b = rng.uniform(0.3, 1.0, 25)
  b /= b.sum()
  eps = 0.2
  out = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, threshold=1e-12, inner_iterations=1,
                                max_iterations=20000)
  assert out.converged
  cost = oracle.sqeuclidean_cost(x, y)
  p = oracle.transport_from_potentials(cost, out.f, out.g, eps)
  q = a[:, None] * b[None, :]
  kl = float((p * np.log(p / q) - p + q).sum())
  primal = oracle.transport_cost_at_cost(cost, out.f, out.g, eps, cost)
  assert math.isclose(primal + eps * kl, out.reg_ot_cost, rel_tol=1e-9)
  np.testing.assert_allclose(p.sum(1), a, rtol=1e-9)
  np.testing.assert_allclose(p.sum(0), b, rtol=1e-9)
  v = rng.normal(size=(2, 30))
  np.testing.assert_allclose(oracle.apply_transport_from_potentials(cost, out.f, out.g, v, eps, axis=0),
                             v @ p, rtol=1e-12)
  w = rng.normal(size=25)
  np.testing.assert_allclose(oracle.apply_transport_from_potentials(cost, out.f, out.g, w, eps, axis=1),
                             p @ w, rtol=1e-12)


def test_oracle_epsilon_schedule():
  """Iteration t of a scheduled solve runs at max(init decay^t, 1) eps
  (epsilon_scheduler.py:47-60) and errors are taken at eps: one iteration
  equals an unscheduled iteration at 4 eps, two equal that followed by one at
  2 eps, and once the schedule reaches eps the solve ends at the unscheduled
  fixed point."""
  x, y, a, b = _problem(40, 30, 3, seed=11)
  cost = oracle.sqeuclidean_cost(x, y)
  ops = oracle.DenseLSE(cost)
  eps = 0.05
  kw = dict(inner_iterations=1, threshold=0.0)
  one = oracle.sinkhorn(ops, a, b, eps, min_iterations=1, max_iterations=1, eps_init=4.0, eps_decay=0.5, **kw)
  big = oracle.sinkhorn(ops, a, b, 4 * eps, min_iterations=1, max_iterations=1, **kw)
  np.testing.assert_array_equal(one.f, big.f)
  np.testing.assert_array_equal(one.g, big.g)
  want_err = oracle.solution_error(ops, one.f, one.g, a, b, eps, 1.0, 1.0, (1,), False)[0]
  assert one.errors[0] == want_err
  assert one.errors[0] != big.errors[0]
  two = oracle.sinkhorn(ops, a, b, eps, min_iterations=2, max_iterations=2, eps_init=4.0, eps_decay=0.5, **kw)
  step = oracle.sinkhorn(ops, a, b, 2 * eps, min_iterations=1, max_iterations=1, init=(big.f, big.g), **kw)
  np.testing.assert_allclose(two.f, step.f, rtol=0, atol=1e-14)
  np.testing.assert_allclose(two.g, step.g, rtol=0, atol=1e-14)
  sched = oracle.sinkhorn(ops, a, b, eps, threshold=1e-10, inner_iterations=5, max_iterations=5000,
                          eps_init=20.0, eps_decay=0.7)
  plain = oracle.sinkhorn(ops, a, b, eps, threshold=1e-10, inner_iterations=5, max_iterations=5000)
  assert sched.converged and plain.converged
  assert math.isclose(sched.reg_ot_cost, plain.reg_ot_cost, rel_tol=1e-9)
  np.testing.assert_allclose(sched.f - sched.f[0], plain.f - plain.f[0], atol=1e-7)


def test_oracle_adaptive_momentum():
  """Momentum(start > 0) (acceleration.py:138-160): Lehmann's weight for a
  known error ratio; a threshold that never holds keeps the constant value;
  otherwise the run equals the constant one up to ``start`` and, from there,
  a constant-momentum run restarted with the weight taken from the errors of
  blocks start // inner - 2 and - 1."""
  w = oracle.momentum_weight(np.array([0.4, 0.1, -1.0]), 6, 1.0, 6, np.inf, 3)
  assert math.isclose(w, 2.0 / (1.0 + math.sqrt(1.0 - 0.25 ** (1.0 / 3.0))), rel_tol=1e-6)
  assert oracle.momentum_weight(np.array([0.4, 0.1]), 5, 1.0, 6, np.inf, 3) == 1.0
  assert oracle.momentum_weight(np.array([0.4, 0.39999]), 6, 1.0, 6, np.inf, 3) == pytest.approx(
      2.0 / (1.0 + math.sqrt(1.0 - 0.99 ** (1.0 / 3.0))), rel=1e-6)  # ratio capped at 0.99
  x, y, a, b = _problem(40, 30, 3, seed=5)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  eps, inner, start = 0.02, 2, 6
  kw = dict(threshold=1e-13, inner_iterations=inner, max_iterations=40)
  const = oracle.sinkhorn(ops, a, b, eps, **kw)
  never = oracle.sinkhorn(ops, a, b, eps, momentum_start=start, momentum_error_threshold=-1.0, **kw)
  np.testing.assert_array_equal(never.f, const.f)
  adapt = oracle.sinkhorn(ops, a, b, eps, momentum_start=start, **kw)
  head = oracle.sinkhorn(ops, a, b, eps, **dict(kw, max_iterations=start))
  np.testing.assert_array_equal(adapt.errors[:start // inner], head.errors)
  e1, e2 = np.float32(head.errors[start // inner - 1]), np.float32(head.errors[start // inner - 2])
  w = 2.0 / (1.0 + math.sqrt(1.0 - float(min(e1 / e2, np.float32(0.99))) ** (1.0 / inner)))
  assert 1.0 < w < 2.0
  tail = oracle.sinkhorn(ops, a, b, eps, momentum=w, init=(head.f, head.g),
                         **dict(kw, max_iterations=40 - start))
  np.testing.assert_allclose(adapt.f, tail.f, rtol=0, atol=1e-9)
  np.testing.assert_allclose(adapt.errors[start // inner:], tail.errors, rtol=1e-5)
  assert adapt.errors[-1] < const.errors[-1]


def test_oracle_reg_ot_cost_grad_against_finite_differences():
  """The gradient at fixed potentials equals the derivative of the converged
  reg_ot_cost (envelope theorem): central differences of full solves in the
  points and, keeping the total mass, in the weights."""
  x, y, a, b = _problem(7, 5, 2, seed=21)
  eps, scale = 0.3, 0.7
  kw = dict(epsilon=eps, scale_cost=1.0 / scale, threshold=1e-14, inner_iterations=1, max_iterations=20000)
  out = oracle.solve_pointcloud(x, y, a, b, **kw)
  assert out.converged
  gx, gy, ga, gb = oracle.reg_ot_cost_grad(x, y, a, b, out.f, out.g, eps, scale)
  h = 1e-5

  def cost(x_, y_, a_, b_):
    return oracle.solve_pointcloud(x_, y_, a_, b_, **kw).reg_ot_cost

  for (i, k) in [(0, 0), (3, 1), (6, 0)]:
    xp, xm = x.copy(), x.copy()
    xp[i, k] += h
    xm[i, k] -= h
    assert abs((cost(xp, y, a, b) - cost(xm, y, a, b)) / (2 * h) - gx[i, k]) < 1e-7
  for (j, k) in [(1, 1), (4, 0)]:
    yp, ym = y.copy(), y.copy()
    yp[j, k] += h
    ym[j, k] -= h
    assert abs((cost(x, yp, a, b) - cost(x, ym, a, b)) / (2 * h) - gy[j, k]) < 1e-7
  ap, am = a.copy(), a.copy()
  ap[2] += h
  ap[5] -= h
  am[2] -= h
  am[5] += h
  assert abs((cost(x, y, ap, b) - cost(x, y, am, b)) / (2 * h) - (ga[2] - ga[5])) < 1e-6
  bp, bm = b.copy(), b.copy()
  bp[0] += h
  bp[3] -= h
  bm[0] -= h
  bm[3] += h
  assert abs((cost(x, y, a, bp) - cost(x, y, a, bm)) / (2 * h) - (gb[0] - gb[3])) < 1e-6


def test_soft_c_transform_and_barycentric_map():
  """soft_c_transform (linear_problem.py:149-171) and barycentric_map
  (potentials.py:52-84): a one-point support's closed form c(x, y) - g -
  eps log b, the hard c-transform -max_j (g_j - C_ij) as eps -> 0, the
  broadcasting of the new points' uniform weights on axis 0 (+ eps log k),
  and the map against central differences of f: T(x) = x - grad f(x) / 2."""
  rng = np.random.default_rng(8)
  x = rng.normal(size=(5, 3))
  y = rng.normal(size=(7, 3))
  g = rng.normal(size=7)
  b = rng.uniform(0.1, 1.0, 7)
  cost = oracle.sqeuclidean_cost(x, y)
  one = oracle.soft_c_transform(cost[:, :1], g[:1], b[:1], 0.3, axis=1)
  np.testing.assert_allclose(one, cost[:, 0] - g[0] - 0.3 * np.log(b[0]), rtol=1e-13)
  hard = -np.max(g[None, :] - cost, axis=1)
  np.testing.assert_allclose(oracle.soft_c_transform(cost, g, b, 1e-4, axis=1), hard, atol=2e-3)
  f = rng.normal(size=5)
  k = cost.shape[1]
  ga = oracle.soft_c_transform(cost, f, np.full(k, 1.0 / k), 0.3, axis=0)
  brute = np.array([-0.3 * np.log(np.sum(np.exp((f - cost[:, j]) / 0.3)) / k) for j in range(k)])
  np.testing.assert_allclose(ga, brute, rtol=1e-12)

  def f_of(pt):
    return oracle.soft_c_transform(oracle.sqeuclidean_cost(pt[None, :], y), g, b, 0.3, axis=1)[0]

  t = oracle.barycentric_map(oracle.sqeuclidean_cost(x, y), g, b, 0.3, y)
  h = 1e-6
  for i in range(len(x)):
    grad = np.array([(f_of(x[i] + h * e) - f_of(x[i] - h * e)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(t[i], x[i] - grad / 2, atol=1e-7)


def test_median_cost_against_sorted_entries():
  """median_cost = the middle entry of the sorted cost matrix (odd n m) or the
  mean of the two middle ones (even n m), on brute-force costs."""
  rng = np.random.default_rng(31)
  for n, m in ((7, 5), (6, 5), (1, 1), (1, 2), (40, 33)):
    x = rng.normal(size=(n, 3))
    y = rng.normal(size=(m, 3))
    c = np.sort(np.array([[np.sum((xi - yj) ** 2) for yj in y] for xi in x]).ravel())
    k = c.size
    want = c[k // 2] if k % 2 else 0.5 * (c[k // 2 - 1] + c[k // 2])
    assert math.isclose(oracle.median_cost(x, y), want, rel_tol=1e-12)
    assert math.isclose(oracle.median_cost(x, y, 0.5), 0.5 * want, rel_tol=1e-12)


@pytest.mark.parametrize("cost_fn", [("pnormp", 1.0), ("pnormp", 1.5), ("pnormp", 2.0), ("sqpnorm", 1.0),
                                     ("sqpnorm", 3.0), ("euclideanp", 1.0), ("euclideanp", 3.0), ("euclidean",)])
def test_ti_cost_against_textbook_norms(cost_fn):
  """ti_cost entry by entry against numpy.linalg.norm of x_i - y_j with the
  costs' definitions (the p-norm, its p-th power over p, half its square, the
  Euclidean norm and its powers), and the special cases that fix the
  constants: PNormP(2) = SqPNorm(2) = ||z||^2 / 2, EuclideanP(2) = ||z||^2,
  PNormP(1) = SqPNorm(1)^(1/2) 2^(1/2) = the l1 norm."""
  rng = np.random.default_rng(7)
  x = rng.normal(size=(6, 4))
  y = rng.normal(size=(5, 4))
  c = oracle.ti_cost(x, y, cost_fn, scale=0.7)
  kind = cost_fn[0]
  for i in range(6):
    for j in range(5):
      z = x[i] - y[j]
      if kind == "pnormp":
        want = np.linalg.norm(z, ord=cost_fn[1]) ** cost_fn[1] / cost_fn[1]
      elif kind == "sqpnorm":
        want = 0.5 * np.linalg.norm(z, ord=cost_fn[1]) ** 2
      elif kind == "euclideanp":
        want = np.linalg.norm(z) ** cost_fn[1]
      else:
        want = np.linalg.norm(z)
      assert math.isclose(c[i, j], 0.7 * want, rel_tol=1e-12)
  sq = oracle.sqeuclidean_cost(x, y)
  np.testing.assert_allclose(oracle.ti_cost(x, y, ("pnormp", 2.0)), 0.5 * sq, rtol=1e-12)
  np.testing.assert_allclose(oracle.ti_cost(x, y, ("sqpnorm", 2.0)), 0.5 * sq, rtol=1e-12)
  np.testing.assert_allclose(oracle.ti_cost(x, y, ("euclideanp", 2.0)), sq, rtol=1e-12)
  l1 = np.abs(x[:, None] - y[None]).sum(-1)
  np.testing.assert_allclose(oracle.ti_cost(x, y, ("pnormp", 1.0)), l1, rtol=1e-12)
  np.testing.assert_allclose(np.sqrt(2.0 * oracle.ti_cost(x, y, ("sqpnorm", 1.0))), l1, rtol=1e-12)


@pytest.mark.parametrize("taus", [(1.0, 1.0), (0.9, 0.8)])
def test_oracle_parallel_updates_general(taus):
  """Parallel updates on x against y (sinkhorn.py:840-846): one iteration
  from (f0, g0) gives g1 = tau_b (eps log b - eps LSE_i((f0_i - C_ij) / eps)) and
  f1 from g0 (not g1), written with scipy's logsumexp; the recorded error is
  |P 1 - a'|_1 + |P^T 1 - b'|_1 of the state (sinkhorn.py:160-189, a' = a or
  a exp(-f / rho) for tau < 1) from the dense coupling; and the converged
  parallel iteration reaches the sequential one's optimum (the same coupling)."""
  ta, tb = taus
  rng = np.random.default_rng(12)
  x = rng.normal(size=(30, 3))
  y = rng.normal(size=(25, 3)) + 0.3
  a = rng.uniform(0.2, 1.0, size=30)
  b = rng.uniform(0.2, 1.0, size=25)
  a /= a.sum()
  b /= b.sum()
  eps = 0.2
  cost = ((x[:, None, :] - y[None]) ** 2).sum(-1)
  f0 = rng.normal(size=30) * 0.1
  g0 = rng.normal(size=25) * 0.1
  one = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, tau_a=ta, tau_b=tb, parallel_dual_updates=True,
                                threshold=0.0, inner_iterations=1, max_iterations=1, init=(f0, g0))
  g1 = tb * (eps * np.log(b) - eps * special.logsumexp((f0[:, None] - cost) / eps, axis=0))
  f1 = ta * (eps * np.log(a) - eps * special.logsumexp((g0[None, :] - cost) / eps, axis=1))
  np.testing.assert_allclose(one.g, g1, rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(one.f, f1, rtol=1e-12, atol=1e-12)
  plan = np.exp((one.f[:, None] + one.g[None] - cost) / eps)
  rho = lambda t: eps * t / (1.0 - t)
  ta_t = a if ta == 1.0 else a * np.exp(-one.f / rho(ta))
  tb_t = b if tb == 1.0 else b * np.exp(-one.g / rho(tb))
  want = np.abs(plan.sum(1) - ta_t).sum() + np.abs(plan.sum(0) - tb_t).sum()
  np.testing.assert_allclose(one.errors[0], want, rtol=1e-10)
  kw = dict(epsilon=eps, tau_a=ta, tau_b=tb, threshold=1e-10, inner_iterations=1, max_iterations=20000)
  # balanced parallel updates oscillate in the gauge without momentum < 1
  par = oracle.solve_pointcloud(x, y, a, b, parallel_dual_updates=True, momentum=0.5, **kw)
  seq = oracle.solve_pointcloud(x, y, a, b, **kw)
  assert par.converged and seq.converged
  p_par = np.exp((par.f[:, None] + par.g[None] - cost) / eps)
  p_seq = np.exp((seq.f[:, None] + seq.g[None] - cost) / eps)
  np.testing.assert_allclose(p_par, p_seq, rtol=1e-6, atol=1e-12)
  assert math.isclose(par.reg_ot_cost, seq.reg_ot_cost, rel_tol=1e-8)
