"""Pins the oracle's unbalanced iteration and cost, plus ent_reg_cost and
dual_cost, by checks that never reuse the formulas they test.

What the reference computes when tau < 1 (sinkhorn.py:803-857, 224-277):

- The iteration g <- tau_b (eps log b - eps LSE_i((f_i - C_ij) / eps)) (and
  its f twin) converges to the optimum of the unbalanced problem whose
  entropy is taken against the counting measure,

    min_P <P, C> + eps sum(P log P - P + 1) + rho_a KL(P 1 | a) + rho_b KL(P^T 1 | b),

  KL(p | q) = sum p log(p / q) - p + q, rho = eps tau / (1 - tau).
- reg_ot_cost is the Lagrange dual function of the problem whose entropy is
  taken against a b^T,

    D(h, k) = inf_{P, r, c >= 0} <P, C> + eps KL(P | a b^T) + rho_a KL(r | a) + rho_b KL(c | b)
                                 + <h, r - P 1> + <k, c - P^T 1>,

  evaluated at h = f - eps log a, k = g - eps log b (a side with tau = 1 holds
  r = a, or c = b, as a constraint).  Its closed form -sum a phi*(-h, rho) -
  ... + eps (sum a sum b - sum P) uses phi*(h, rho) = rho (exp(h / rho) - 1),
  the convex conjugate of rho KL (unbalanced_functions.py:19-21).

The two problems differ when tau < 1, so reg_ot_cost is not the primal value
of the returned coupling there (it is in the balanced case).  The build
reproduces the reference's numbers.  The tests below pin the fixed point by a
direct L-BFGS minimisation of the first problem and the cost by a numerical
infimum of the Lagrangian (no phi*, no closed form), at converged and at
arbitrary potentials; planted slips (phi*'s sign, rho against 1 / rho, f in
place of f - eps log a, the total-mass term) fail them.
"""

import functools
import math

import numpy as np
import pytest
from scipy import optimize, special

from oracle import sinkhorn64 as oracle


def _kl(p, q):
  return float(np.sum(special.rel_entr(p, q) - p + q))


def _rho(eps, tau):
  return math.inf if tau == 1.0 else eps * tau / (1.0 - tau)


def _clouds(n, m, d, seed, mass_b=1.0):
  rng = np.random.default_rng(seed)
  x = rng.normal(size=(n, d)) * 0.6
  y = rng.normal(size=(m, d)) * 0.6 + 0.3
  a = rng.uniform(0.3, 1.0, size=n)
  b = rng.uniform(0.3, 1.0, size=m)
  return x, y, a / a.sum(), mass_b * b / b.sum()


def _solve(x, y, a, b, eps, ta, tb):
  out = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, tau_a=ta, tau_b=tb, threshold=1e-13,
                                inner_iterations=1, max_iterations=50000)
  assert out.converged
  return out


def _lbfgs(fun, x0):
  res = optimize.minimize(fun, x0, jac=True, method="L-BFGS-B",
                          options=dict(maxiter=50000, ftol=1e-16, gtol=1e-13, maxcor=50))
  return res


def _lagrangian_dual(cost, a, b, eps, ta, tb, h, k):
  """inf over P, r, c of the a b^T-referenced Lagrangian, by L-BFGS over
  log P, log r, log c (the objective is separable and convex in P, r, c)."""
  n, m = cost.shape
  q = np.outer(a, b)
  ra, rb = _rho(eps, ta), _rho(eps, tb)
  fin_h = np.where(np.isfinite(h), h, 0.0)
  fin_k = np.where(np.isfinite(k), k, 0.0)

  def fun(z):
    u = z[:n * m].reshape(n, m)
    p = np.exp(u)
    lin = cost - h[:, None] - k[None, :]
    val = np.sum(p * lin) + eps * np.sum(p * (u - np.log(q)) - p + q)
    gu = p * (lin + eps * (u - np.log(q)))
    grads = [gu.ravel()]
    off = n * m
    if ta < 1.0:
      t = z[off:off + n]
      r = np.exp(t)
      val += ra * np.sum(r * (t - np.log(a)) - r + a) + np.sum(fin_h * r)
      grads.append(r * (ra * (t - np.log(a)) + fin_h))
      off += n
    else:
      val += np.sum(fin_h * a)
    if tb < 1.0:
      t = z[off:off + m]
      c = np.exp(t)
      val += rb * np.sum(c * (t - np.log(b)) - c + b) + np.sum(fin_k * c)
      grads.append(c * (rb * (t - np.log(b)) + fin_k))
    else:
      val += np.sum(fin_k * b)
    return float(val), np.concatenate(grads)

  z0 = [np.log(q).ravel()]
  if ta < 1.0:
    z0.append(np.log(a))
  if tb < 1.0:
    z0.append(np.log(b))
  return _lbfgs(fun, np.concatenate(z0)).fun


TAUS = [(0.8, 0.9), (1.0, 0.9), (0.9, 1.0), (0.95, 0.95), (0.5, 0.7)]


@pytest.mark.parametrize("ta,tb", [(0.8, 0.9), (0.95, 0.95), (0.5, 0.7)])
@pytest.mark.parametrize("mass_b", [1.0, 1.4])
def test_unbalanced_fixed_point_is_the_counting_measure_optimum(ta, tb, mass_b):
  """The converged coupling P = exp((f + g - C) / eps) attains the minimum of
  the counting-measure-referenced primal found by L-BFGS over log P, and the
  problem's dual at (f, g) has the same value (strong duality)."""
  x, y, a, b = _clouds(7, 9, 2, seed=3, mass_b=mass_b)
  eps = 0.2
  out = _solve(x, y, a, b, eps, ta, tb)
  cost = oracle.sqeuclidean_cost(x, y)
  ra, rb = _rho(eps, ta), _rho(eps, tb)
  n, m = cost.shape

  def primal(u):
    u = u.reshape(n, m)
    p = np.exp(u)
    r, c = p.sum(1), p.sum(0)
    val = np.sum(p * cost) + eps * np.sum(p * u - p + 1.0) + ra * _kl(r, a) + rb * _kl(c, b)
    grad = cost + eps * u + ra * np.log(r / a)[:, None] + rb * np.log(c / b)[None, :]
    return float(val), (p * grad).ravel()

  res = _lbfgs(primal, np.log(np.outer(a, b)).ravel())
  p_star = oracle.transport_from_potentials(cost, out.f, out.g, eps)
  at_star = primal(np.log(p_star).ravel())[0]
  # L-BFGS stops just above the minimum; the fixed point is at least as low
  assert math.isclose(at_star, res.fun, rel_tol=1e-9), (at_star, res.fun)
  assert at_star <= res.fun + 1e-12 * abs(res.fun)
  np.testing.assert_allclose(np.exp(res.x).reshape(n, m), p_star, rtol=1e-4, atol=1e-7)
  dual = (-ra * np.sum(a * (np.exp(-out.f / ra) - 1.0)) - rb * np.sum(b * (np.exp(-out.g / rb) - 1.0))
          - eps * np.sum(p_star - 1.0))
  assert math.isclose(dual, at_star, rel_tol=1e-10)


@pytest.mark.parametrize("ta,tb", TAUS)
@pytest.mark.parametrize("where", ["converged", "arbitrary"])
def test_reg_ot_cost_is_the_lagrangian_dual_function(ta, tb, where):
  """compute_kl_reg_cost(f, g) = D(f - eps log a, g - eps log b), the
  infimum of the a b^T-referenced Lagrangian, at the solver's potentials and
  at arbitrary ones, with unequal masses."""
  x, y, a, b = _clouds(6, 8, 2, seed=5, mass_b=1.3)
  eps = 0.25
  cost = oracle.sqeuclidean_cost(x, y)
  ops = oracle.DenseLSE(cost)
  if where == "converged":
    out = _solve(x, y, a, b, eps, ta, tb)
    f, g = out.f, out.g
  else:
    rng = np.random.default_rng(17)
    f = rng.normal(size=6) * 0.3
    g = rng.normal(size=8) * 0.3
  got = oracle.compute_kl_reg_cost(ops, f, g, a, b, eps, ta, tb)
  want = _lagrangian_dual(cost, a, b, eps, ta, tb, f - eps * np.log(a), g - eps * np.log(b))
  assert math.isclose(got, want, rel_tol=1e-9, abs_tol=1e-12), (got, want)
  if where == "converged":
    assert got == out.reg_ot_cost


def test_reg_ot_cost_one_point_per_side():
  """n = m = 1: the fixed point in closed form (f = ta (A - g), g = tb (B - f)
  with A = eps log a + C, B = eps log b + C) and the cost as the Lagrangian's
  infimum there."""
  x = np.array([[0.0, 0.0]])
  y = np.array([[0.6, 0.8]])
  a = np.array([0.7])
  b = np.array([1.3])
  eps, ta, tb = 0.2, 0.8, 0.9
  A = eps * math.log(0.7) + 1.0
  B = eps * math.log(1.3) + 1.0
  f_star = ta * (A - tb * B) / (1 - ta * tb)
  g_star = tb * (B - f_star)
  out = _solve(x, y, a, b, eps, ta, tb)
  assert math.isclose(out.f[0], f_star, rel_tol=1e-10) and math.isclose(out.g[0], g_star, rel_tol=1e-10)
  want = _lagrangian_dual(np.array([[1.0]]), a, b, eps, ta, tb, np.array([f_star - eps * math.log(0.7)]),
                          np.array([g_star - eps * math.log(1.3)]))
  assert math.isclose(out.reg_ot_cost, want, rel_tol=1e-10), (out.reg_ot_cost, want)


def test_reg_ot_cost_continuous_at_tau_one():
  """tau -> 1 recovers the balanced problem: potentials and cost of tau = 1 - 1e-7
  agree with the balanced ones (a rho against 1 / rho slip diverges here)."""
  x, y, a, b = _clouds(8, 10, 3, seed=4)
  eps = 0.2
  bal = _solve(x, y, a, b, eps, 1.0, 1.0)
  near = oracle.solve_pointcloud(x, y, a, b, epsilon=eps, tau_a=1 - 1e-7, tau_b=1 - 1e-7, threshold=1e-9,
                                 inner_iterations=10, max_iterations=50000)
  assert math.isclose(near.reg_ot_cost, bal.reg_ot_cost, rel_tol=1e-5)
  gauge = np.mean(near.f - bal.f)
  np.testing.assert_allclose(near.f - gauge, bal.f, atol=1e-4)


def test_ent_reg_cost_and_dual_cost_balanced():
  """Balanced, at convergence with sum a = sum b = 1: ent_reg_cost (sinkhorn.py:350-370)
  and dual_cost (:337-343) both equal <P, C> - eps H(P), H(P) = -sum P log P,
  computed from the coupling; and the duality gap <P, C> - dual_cost is
  positive and vanishes with eps, as the reference's primal/dual test
  (sinkhorn_test.py:595-614) checks."""
  x, y, a, b = _clouds(10, 14, 2, seed=9)
  eps = 0.1
  out = _solve(x, y, a, b, eps, 1.0, 1.0)
  cost = oracle.sqeuclidean_cost(x, y)
  p = oracle.transport_from_potentials(cost, out.f, out.g, eps)
  want = float(np.sum(p * cost) + eps * np.sum(special.xlogy(p, p)))
  ent = oracle.ent_reg_cost(out.reg_ot_cost, a, b, eps)
  dual = oracle.dual_cost(out.f, out.g, a, b)
  assert math.isclose(ent, want, rel_tol=1e-9)
  assert math.isclose(dual, want, rel_tol=1e-9)
  gaps = []
  for e in (0.1, 0.03, 0.01):
    o = oracle.solve_pointcloud(x, y, a, b, epsilon=e, threshold=1e-10, inner_iterations=10,
                                max_iterations=200000)
    pe = oracle.transport_from_potentials(cost, o.f, o.g, e)
    gaps.append(float(np.sum(pe * cost)) - oracle.dual_cost(o.f, o.g, a, b))
  assert all(g > 0 for g in gaps) and gaps[0] > gaps[1] > gaps[2]
  # zero weights: their -inf potentials do not enter dual_cost
  a0 = a.copy()
  a0[2] = 0.0
  a0 /= a0.sum()
  o = _solve(x, y, a0, b, eps, 1.0, 1.0)
  assert np.isneginf(o.f[2]) and np.isfinite(oracle.dual_cost(o.f, o.g, a0, b))


# --------------------------------------------------------------------------
# The checks above fail on plausible slips planted in the oracle
# --------------------------------------------------------------------------
def _slip_case():
  x, y, a, b = _clouds(6, 8, 2, seed=5, mass_b=1.3)
  rng = np.random.default_rng(17)
  return oracle.sqeuclidean_cost(x, y), a, b, 0.25, rng.normal(size=6) * 0.3, rng.normal(size=8) * 0.3


@functools.lru_cache(maxsize=None)
def _slip_want(ta=0.8, tb=0.9):
  cost, a, b, eps, f, g = _slip_case()
  return _lagrangian_dual(cost, a, b, eps, ta, tb, f - eps * np.log(a), g - eps * np.log(b))


def _dual_gap(ta=0.8, tb=0.9):
  cost, a, b, eps, f, g = _slip_case()
  got = oracle.compute_kl_reg_cost(oracle.DenseLSE(cost), f, g, a, b, eps, ta, tb)
  want = _slip_want(ta, tb)
  return abs(got - want) / abs(want)


def _with_rho(rho_fn, fn, *args, **kw):
  saved = oracle.rho
  oracle.rho = rho_fn
  try:
    return fn(*args, **kw)
  finally:
    oracle.rho = saved


@pytest.mark.parametrize("slip", ["phi_star_sign", "phi_star_inv_rho", "rho_inverted", "no_log_weight",
                                  "no_total_mass"])
def test_planted_slips_in_the_unbalanced_cost_fail(monkeypatch, slip):
  assert _dual_gap() < 1e-9
  phi, rho, kl = oracle.phi_star, oracle.rho, oracle.compute_kl_reg_cost
  if slip == "phi_star_sign":
    monkeypatch.setattr(oracle, "phi_star", lambda h, r: r * (np.exp(-h / r) - 1.0))
  elif slip == "phi_star_inv_rho":
    monkeypatch.setattr(oracle, "phi_star", lambda h, r: (np.exp(h * r) - 1.0) / r)
  elif slip == "rho_inverted":
    monkeypatch.setattr(oracle, "compute_kl_reg_cost",
                        lambda *args, **kw: _with_rho(lambda e, t: (1.0 - t) / (e * t), kl, *args, **kw))
  elif slip == "no_log_weight":
    def no_log(ops, f, g, a, b, eps, tau_a=1.0, tau_b=1.0):
      div_a = -np.sum(np.where(a > 0, a * phi(-f, rho(eps, tau_a)), 0.0))
      div_b = -np.sum(np.where(b > 0, b * phi(-g, rho(eps, tau_b)), 0.0))
      total = np.sum(oracle.marginal_from_potentials(ops, f, g, eps, 0))
      return div_a + div_b + eps * (np.sum(a) * np.sum(b) - total)
    monkeypatch.setattr(oracle, "compute_kl_reg_cost", no_log)
  else:
    def dropped(ops, f, g, a, b, eps, tau_a=1.0, tau_b=1.0):
      total = np.sum(oracle.marginal_from_potentials(ops, f, g, eps, 0))
      return kl(ops, f, g, a, b, eps, tau_a, tau_b) - eps * (np.sum(a) * np.sum(b) - total)
    monkeypatch.setattr(oracle, "compute_kl_reg_cost", dropped)
  assert _dual_gap() > 1e-4


# --------------------------------------------------------------------------
# Recentering (sinkhorn.py:76-114, 803-857) with both taus < 1.
#
# Along a translation (f + l, g - l) the coupling does not change and the
# marginal terms of the unbalanced dual, -rho_a sum a exp(-(f + l) / rho_a) -
# rho_b sum b exp(-(g - l) / rho_b), are maximised where
# exp(-l / rho_a) A = exp(l / rho_b) B (A, B the sums at l = 0).  At the
# optimum of the dual that fixes l = 0, so the shift that ``recenter`` applies
# to any translate of the optimum must map it back exactly; and the
# translation-invariant iteration (the k / xi / smin terms) must converge to
# the same potentials as the plain unbalanced iteration, whose fixed point is
# pinned above by a direct minimisation of the primal.
# --------------------------------------------------------------------------
RC_TAUS = [(0.8, 0.9), (0.95, 0.95), (0.6, 0.85)]


@functools.lru_cache(maxsize=None)
def _rc_case(ta, tb):
  x, y, a, b = _clouds(10, 7, 3, seed=41, mass_b=1.2)
  eps = 0.4
  return x, y, a, b, eps, _solve(x, y, a, b, eps, ta, tb)


@pytest.mark.parametrize("ta,tb", RC_TAUS)
@pytest.mark.parametrize("lam", [0.0, 0.7, -1.3])
def test_recenter_maps_translates_of_the_optimum_back(ta, tb, lam):
  x, y, a, b, eps, ref = _rc_case(ta, tb)
  f, g = oracle.recenter(ref.f + lam, ref.g - lam, a, b, eps, ta, tb)
  np.testing.assert_allclose(f, ref.f, rtol=0, atol=1e-9)
  np.testing.assert_allclose(g, ref.g, rtol=0, atol=1e-9)


def _rc_solve(x, y, a, b, eps, ta, tb, max_iterations=20000):
  return oracle.solve_pointcloud(x, y, a, b, epsilon=eps, tau_a=ta, tau_b=tb, threshold=1e-12,
                                 inner_iterations=1, max_iterations=max_iterations,
                                 recenter_potentials=True)


@pytest.mark.parametrize("ta,tb", RC_TAUS)
def test_translation_invariant_iteration_reaches_the_unbalanced_optimum(ta, tb):
  x, y, a, b, eps, ref = _rc_case(ta, tb)
  out = _rc_solve(x, y, a, b, eps, ta, tb)
  assert out.converged
  np.testing.assert_allclose(out.f, ref.f, rtol=0, atol=1e-8)
  np.testing.assert_allclose(out.g, ref.g, rtol=0, atol=1e-8)
  assert abs(out.reg_ot_cost - ref.reg_ot_cost) <= 1e-9 * abs(ref.reg_ot_cost)


def _tiu_g(ops, a, b, eps, ta, tb, f, g):
  new_g = tb * oracle.update_potential(ops, f, g, np.log(b), eps, axis=0)
  new_g = new_g - oracle.recenter_k(tb, tb, ta, tb) * oracle.smin(f, a, eps, ta)
  return new_g + oracle.recenter_xi(tb, ta, ta, tb) * oracle.smin(new_g, b, eps, tb)


def _tiu_f(ops, a, b, eps, ta, tb, f, g):
  new_f = ta * oracle.update_potential(ops, f, g, np.log(a), eps, axis=1)
  new_f = new_f - oracle.recenter_k(ta, ta, ta, tb) * oracle.smin(g, b, eps, tb)
  return new_f + oracle.recenter_xi(ta, tb, ta, tb) * oracle.smin(new_f, a, eps, ta)


def _identity_residual(ta, tb, seed=3):
  """max |g_new - tau_b U(f) - (1 - tau_b) s| and its f twin after one
  recentred iteration of the oracle from random potentials (momentum 1, so
  the iteration's potentials are the updates themselves)."""
  x, y, a, b, eps, _ = _rc_case(ta, tb)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  rng = np.random.default_rng(seed)
  f0, g0 = rng.normal(size=len(a)), rng.normal(size=len(b))
  ra, rb = _rho(eps, ta), _rho(eps, tb)
  tau = ra * rb / (ra + rb)

  def shift(f, g):
    return tau * (special.logsumexp(-f / ra, b=a) - special.logsumexp(-g / rb, b=b))

  # the oracle's iteration, read through its two halves
  g1 = _tiu_g(ops, a, b, eps, ta, tb, f0, g0)
  f1 = _tiu_f(ops, a, b, eps, ta, tb, f0, g1)
  out = oracle.sinkhorn(ops, a, b, eps, tau_a=ta, tau_b=tb, init=(f0, g0), inner_iterations=1,
                        min_iterations=1, max_iterations=1, recenter_potentials=True)
  # the iteration's state is (f1, g1); its output is that state recentred
  s1 = shift(f1, g1)
  np.testing.assert_allclose(out.f, f1 + s1, rtol=0, atol=1e-12)
  np.testing.assert_allclose(out.g, g1 - s1, rtol=0, atol=1e-12)
  ug = tb * oracle.update_potential(ops, f0, g0, np.log(b), eps, axis=0)
  uf = ta * oracle.update_potential(ops, f0, g1, np.log(a), eps, axis=1)
  rg = np.max(np.abs(g1 - (ug + (1.0 - tb) * shift(f0, g1))))
  rf = np.max(np.abs(f1 - (uf - (1.0 - ta) * shift(f1, g1))))
  return rg, rf


@pytest.mark.parametrize("ta,tb", RC_TAUS)
def test_recentred_updates_are_translation_optimal(ta, tb):
  rg, rf = _identity_residual(ta, tb)
  assert rg < 1e-12 and rf < 1e-12


@pytest.mark.parametrize("slip", ["k_sign", "xi_swapped", "xi_is_k", "smin_sign"])
def test_planted_slips_in_the_recentering_terms_fail(monkeypatch, slip):
  ta, tb = 0.8, 0.9
  k, xi, sm = oracle.recenter_k, oracle.recenter_xi, oracle.smin
  if slip == "k_sign":
    monkeypatch.setattr(oracle, "recenter_k", lambda ti, tj, t1, t2: -k(ti, tj, t1, t2))
  elif slip == "xi_swapped":
    monkeypatch.setattr(oracle, "recenter_xi", lambda ti, tj, t1, t2: xi(tj, ti, t1, t2))
  elif slip == "xi_is_k":
    monkeypatch.setattr(oracle, "recenter_xi", lambda ti, tj, t1, t2: k(ti, tj, t1, t2))
  else:
    monkeypatch.setattr(oracle, "smin", lambda p, w, e, t: -sm(p, w, e, t))
  rg, rf = _identity_residual(ta, tb)
  assert max(rg, rf) > 1e-6
