"""GPU parity of the HIP path against the FP64 oracle.

Every test calls the C ABI through the ``ott_amd`` binding on seeded inputs and
compares with ``oracle/sinkhorn64.py``.  Tolerances: float32 path rel 1e-4,
bf16-split path rel 1e-3 (on potentials, relative to their scale, and on
reg_ot_cost).
"""

import math

import numpy as np
import pytest
import torch

from oracle import sinkhorn64 as oracle

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16_split": 1e-3}


def _cloud(n, m, d, seed, shift=0.2, zero_a=None, zero_b=None):
  rng = np.random.default_rng(seed)
  x = (rng.normal(size=(n, d)) / np.sqrt(d)).astype(np.float32)
  y = (rng.normal(size=(m, d)) / np.sqrt(d) + shift).astype(np.float32)
  a = rng.uniform(0.2, 1.2, size=n)
  b = rng.uniform(0.2, 1.2, size=m)
  if zero_a is not None:
    a[zero_a] = 0.0
  if zero_b is not None:
    b[zero_b] = 0.0
  return x, y, (a / a.sum()).astype(np.float32), (b / b.sum()).astype(np.float32)


def _geom(x, y, precision, **kw):
  from ott_amd.geometry import pointcloud
  return pointcloud.PointCloud(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                               precision=precision, **kw)


def _close(got, want, rtol, what):
  got = np.asarray(got, dtype=np.float64)
  want = np.asarray(want, dtype=np.float64)
  fin = np.isfinite(want)
  assert np.array_equal(fin, np.isfinite(got)), f"{what}: finiteness differs"
  scale = max(1.0, float(np.max(np.abs(want[fin])))) if fin.any() else 1.0
  err = float(np.max(np.abs(got[fin] - want[fin]))) if fin.any() else 0.0
  assert err <= rtol * scale, f"{what}: max abs err {err:.3e} > {rtol:.1e} * {scale:.3e}"
  if (~fin).any():
    assert np.array_equal(got[~fin], want[~fin]), f"{what}: non-finite entries differ"


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("d", [2, 5, 16, 64, 100])
@pytest.mark.parametrize("axis", [0, 1])
def test_apply_lse_kernel(precision, d, axis):
  x, y, a, b = _cloud(300, 517, d, seed=d, zero_b=[3])
  rng = np.random.default_rng(7)
  f = rng.normal(size=300).astype(np.float32) * 0.1
  g = rng.normal(size=517).astype(np.float32) * 0.1
  g[3] = -np.inf
  eps = 0.07
  geom = _geom(x, y, precision, epsilon=eps)
  got, sign = geom.apply_lse_kernel(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(),
                                    eps, axis=axis)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  want = oracle.apply_lse_kernel(ops, f.astype(np.float64), g.astype(np.float64), eps, axis)
  _close(got.cpu().numpy(), want, TOL[precision], f"apply_lse axis={axis}")
  assert float(sign[0]) == 1.0


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("axis", [0, 1])
def test_apply_lse_kernel_at_2_24_pairs(precision, axis):
  """apply_lse_kernel (geometry.py:260-299) where the library's default tile
  configuration switches to the large-problem one (2^24 pairs, d = 64)."""
  x, y, a, b = _cloud(4096, 4096, 64, seed=64)
  rng = np.random.default_rng(8)
  f = rng.normal(size=4096).astype(np.float32) * 0.2
  g = rng.normal(size=4096).astype(np.float32) * 0.2
  eps = 0.1
  print(f"apply_lse 4096x4096 d=64 {precision} axis={axis}", flush=True)
  geom = _geom(x, y, precision, epsilon=eps)
  got, _ = geom.apply_lse_kernel(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(), eps, axis=axis)
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  want = oracle.apply_lse_kernel(ops, f.astype(np.float64), g.astype(np.float64), eps, axis)
  _close(got.cpu().numpy(), want, TOL[precision], f"apply_lse 2^24 axis={axis}")


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("axis", [0, 1])
def test_update_potential_and_marginal(precision, axis):
  x, y, a, b = _cloud(257, 129, 32, seed=11, zero_a=[0], zero_b=[5])
  f = np.where(a > 0, 0.0, -np.inf).astype(np.float32)
  g = np.where(b > 0, 0.0, -np.inf).astype(np.float32)
  geom = _geom(x, y, precision, epsilon=0.05, relative_epsilon="mean")
  eps = geom.epsilon
  want_eps = oracle.resolve_epsilon(x, y, 0.05, "mean")
  assert abs(eps - want_eps) <= 1e-6 * want_eps
  ops = oracle.DenseLSE(oracle.sqeuclidean_cost(x, y))
  logw = np.log(b if axis == 0 else a)
  with np.errstate(divide="ignore"):
    got = geom.update_potential(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(),
                                torch.from_numpy(logw.astype(np.float32)).cuda(), axis=axis)
  want = oracle.update_potential(ops, f.astype(np.float64), g.astype(np.float64),
                                 logw, eps, axis)
  _close(got.cpu().numpy(), want, TOL[precision], "update_potential")
  # The marginal after a half-step equals the target exactly (a property of Sinkhorn).
  fg = (f, got.cpu().numpy()) if axis == 0 else (got.cpu().numpy(), g)
  marg = geom.marginal_from_potentials(torch.from_numpy(fg[0]).cuda(),
                                       torch.from_numpy(fg[1]).cuda(), axis=axis)
  target = b if axis == 0 else a
  np.testing.assert_allclose(marg.cpu().numpy(), target, rtol=20 * TOL[precision], atol=1e-7)


def _solve(x, y, a, b, precision, eps_kw, **kw):
  from ott_amd.solvers import linear
  geom = _geom(x, y, precision, **eps_kw)
  at = None if a is None else torch.from_numpy(a).cuda()
  bt = None if b is None else torch.from_numpy(b).cuda()
  tau_a = kw.pop("tau_a", 1.0)
  tau_b = kw.pop("tau_b", 1.0)
  return linear.solve(geom, a=at, b=bt, tau_a=tau_a, tau_b=tau_b, **kw)


def _check_solve(out, ref, precision, same_iters=True):
  tol = TOL[precision]
  assert out.converged == ref.converged
  if getattr(out, "solve_n_iters", None) is not None:
    # the count the solve reported is the one its errors give (sinkhorn.py:431-434)
    assert out.solve_n_iters == int((out.errors != -1).sum()) * out.inner_iterations
  if same_iters:
    assert out.n_iters == ref.n_iters, (out.n_iters, ref.n_iters)
    errs = out.errors.cpu().numpy()
    assert np.array_equal(errs == -1, ref.errors == -1)
    assert np.array_equal(np.isinf(errs), np.isinf(ref.errors))
    run = (ref.errors != -1) & np.isfinite(ref.errors)
    np.testing.assert_allclose(errs[run], ref.errors[run], rtol=20 * tol, atol=1e-6)
    _close(out.f.cpu().numpy(), ref.f, tol, "f")
    _close(out.g.cpu().numpy(), ref.g, tol, "g")
  rel = abs(out.reg_ot_cost - ref.reg_ot_cost) / max(abs(ref.reg_ot_cost), 1e-12)
  assert rel <= tol, f"reg_ot_cost {out.reg_ot_cost} vs {ref.reg_ot_cost} (rel {rel:.2e})"


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_solve_uniform_shifted_d2(precision):
  from ott_amd import datasets
  pb = datasets.uniform_shifted(1024, 1024, 2, seed=0)
  out = _solve(pb.x, pb.y, pb.a, pb.b, precision, dict(epsilon=0.05, relative_epsilon="mean"))
  ref = oracle.solve_pointcloud(pb.x, pb.y, pb.a, pb.b, epsilon=0.05, relative_epsilon="mean")
  _check_solve(out, ref, precision)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_solve_zero_weights(precision, use_graph):
  x, y, a, b = _cloud(173, 291, 4, seed=3, zero_a=[0], zero_b=[3])
  out = _solve(x, y, a, b, precision, dict(epsilon=0.1), use_graph=use_graph)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1)
  _check_solve(out, ref, precision)
  assert np.isneginf(out.f.cpu().numpy()[0]) and np.isneginf(out.g.cpu().numpy()[3])


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("tau", [(0.8, 0.9), (1.0, 0.9), (0.95, 0.95)])
def test_solve_unbalanced(precision, tau):
  x, y, a, b = _cloud(333, 222, 16, seed=5)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.1), tau_a=tau[0], tau_b=tau[1])
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, tau_a=tau[0], tau_b=tau[1])
  _check_solve(out, ref, precision)


def test_solve_min_iterations_and_scan():
  x, y, a, b = _cloud(120, 90, 4, seed=9)
  out = _solve(x, y, a, b, "fp32", dict(epsilon=0.1), min_iterations=34)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, min_iterations=34)
  _check_solve(out, ref, "fp32")
  errs = out.errors.cpu().numpy()
  assert np.isinf(errs[:3]).all() and errs[3] > 0
  out = _solve(x, y, a, b, "fp32", dict(epsilon=0.1), min_iterations=50, max_iterations=50)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, min_iterations=50, max_iterations=50)
  _check_solve(out, ref, "fp32")
  assert out.converged and 0 < float(out.errors[-1]) < 1e-3


@pytest.mark.parametrize("kw", [dict(norm_error=2), dict(inner_iterations=7),
                                dict(momentum_value=1.3), dict(max_iterations=25),
                                dict(threshold=1e-5, max_iterations=40)])
def test_solve_options(kw):
  from ott_amd.solvers.linear import acceleration
  x, y, a, b = _cloud(150, 140, 8, seed=21)
  okw = dict(kw)
  if "momentum_value" in okw:
    v = okw.pop("momentum_value")
    gkw = dict(okw, momentum=acceleration.Momentum(value=v))
    ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, momentum=v, **okw)
  else:
    gkw = okw
    ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, **okw)
  out = _solve(x, y, a, b, "fp32", dict(epsilon=0.1), **gkw)
  _check_solve(out, ref, "fp32")


def test_solve_warm_start_restart():
  """A restart from converged potentials stops after one block (sinkhorn_test.py:491-546)."""
  from ott_amd.geometry import pointcloud
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  x, y, a, b = _cloud(200, 180, 8, seed=4)
  geom = _geom(x, y, "fp32", epsilon=0.05)
  prob = linear_problem.LinearProblem(geom, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
  solver = sinkhorn.Sinkhorn(threshold=1e-4)
  out = solver(prob)
  assert out.converged and out.n_iters > 10
  out2 = solver(prob, init=(out.f, out.g))
  assert out2.converged and out2.n_iters == 10
  rel = abs(out2.reg_ot_cost - out.reg_ot_cost) / abs(out.reg_ot_cost)
  assert rel < 1e-5


class _GivenF:
  """An initializer that predicts f only: the solve takes f as given and sets g
  itself (OTTX_INIT_F_GIVEN), as with the learned warm start."""
  predicts_f = True

  def __init__(self, f):
    self.f = f

  def predict_f(self, ot_prob, out=None):
    out.copy_(self.f)
    return out


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("start", ["converged", "off_sample"])
def test_first_pass_shift_estimates_leave_the_solve_unchanged(precision, start):
  """A start from f alone estimates the first g-update's log-sum-exp from 64
  sampled rows and streams that pass with a fixed shift; a tile whose sums
  leave range re-streams with the running max.  Either way the solve must be
  the one of the same start given as (f, g) explicitly (first g-update from
  f, sinkhorn.py:794-807).  "off_sample" raises f by 100 eps on every row the
  sample misses, so the estimate is far outside the window and the fallback
  runs."""
  from ott_amd.problems.linear import linear_problem
  from ott_amd.solvers.linear import sinkhorn
  n, m = 700, 513
  x, y, a, b = _cloud(n, m, 16, seed=12)
  eps = 0.05
  geom = _geom(x, y, precision, epsilon=eps)
  ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
  solver = sinkhorn.Sinkhorn(threshold=1e-4, inner_iterations=1)
  f0 = solver(linear_problem.LinearProblem(geom, ta, tb)).f.clone()
  if start == "off_sample":
    sampled = (np.arange(64) * n) // 64  # the probe's rows
    off = np.ones(n, dtype=bool)
    off[sampled] = False
    f0[torch.from_numpy(off).cuda()] += 100.0 * eps
  fast = sinkhorn.Sinkhorn(threshold=1e-4, inner_iterations=1, initializer=_GivenF(f0))(
      linear_problem.LinearProblem(_geom(x, y, precision, epsilon=eps), ta, tb))
  explicit = solver(linear_problem.LinearProblem(_geom(x, y, precision, epsilon=eps), ta, tb),
                    init=(f0, torch.zeros(m, device=f0.device)))
  assert fast.converged and explicit.converged
  assert fast.n_iters == explicit.n_iters
  assert abs(fast.reg_ot_cost - explicit.reg_ot_cost) <= 1e-6 * abs(explicit.reg_ot_cost)
  torch.testing.assert_close(fast.f, explicit.f, rtol=1e-5, atol=1e-5 * float(explicit.f.abs().max()))
  torch.testing.assert_close(fast.g, explicit.g, rtol=1e-5, atol=1e-5 * float(explicit.g.abs().max()))


def test_determinism():
  x, y, a, b = _cloud(700, 900, 64, seed=8)
  o1 = _solve(x, y, a, b, "bf16_split", dict(epsilon=0.05, relative_epsilon="mean"))
  o2 = _solve(x, y, a, b, "bf16_split", dict(epsilon=0.05, relative_epsilon="mean"))
  assert torch.equal(o1.f, o2.f) and torch.equal(o1.g, o2.g)
  assert o1.reg_ot_cost == o2.reg_ot_cost


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("masses", ["equal", "unequal"])
def test_recenter_potentials_balanced(precision, masses):
  """recenter_potentials=True (sinkhorn.py:76-114, 1008-1009) on a balanced
  problem: f has zero mean over its finite entries, f + g is that of the
  plain solve, and outputs and reg_ot_cost match the oracle's.  With unequal
  masses (sum b = 1.5 sum a, a fixed number of iterations) the shift moves the
  dual terms of reg_ot_cost by c (sum b - sum a)."""
  x, y, a, b = _cloud(173, 291, 4, seed=5, zero_a=[2], zero_b=[7])
  kw = dict(inner_iterations=1, threshold=1e-4)
  if masses == "unequal":
    b = (b * 1.5).astype(np.float32)
    kw = dict(inner_iterations=1, min_iterations=7, max_iterations=7)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.1), recenter_potentials=True, **kw)
  plain = _solve(x, y, a, b, precision, dict(epsilon=0.1), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, recenter_potentials=True, **kw)
  _check_solve(out, ref, precision)
  f, g = out.f.cpu().numpy().astype(np.float64), out.g.cpu().numpy().astype(np.float64)
  fp, gp = plain.f.cpu().numpy().astype(np.float64), plain.g.cpu().numpy().astype(np.float64)
  fin = np.isfinite(f)
  assert np.array_equal(fin, np.isfinite(fp)) and np.isneginf(f[2]) and np.isneginf(g[7])
  scale = np.abs(fp[fin]).max()
  assert abs(f[fin].mean()) <= 1e-5 * scale
  c = fp[fin].mean()
  np.testing.assert_allclose(f[fin], fp[fin] - c, atol=1e-5 * scale)
  gfin = np.isfinite(g)
  np.testing.assert_allclose(g[gfin], gp[gfin] + c, atol=1e-5 * scale)
  shift = c * (float(b.astype(np.float64).sum()) - float(a.astype(np.float64).sum()))
  assert abs(out.reg_ot_cost - (plain.reg_ot_cost + shift)) <= 1e-5 * max(1.0, abs(plain.reg_ot_cost))
  # the errors do not depend on the gauge
  np.testing.assert_array_equal(out.errors.cpu().numpy(), plain.errors.cpu().numpy())


def test_recenter_potentials_unbalanced_one_side_is_a_no_op():
  """With tau_a = 1 and tau_b < 1 the reference leaves the potentials as they are (sinkhorn.py:104-106)."""
  x, y, a, b = _cloud(150, 120, 3, seed=6)
  kw = dict(inner_iterations=1, tau_b=0.9)
  out = _solve(x, y, a, b, "fp32", dict(epsilon=0.1), recenter_potentials=True, **dict(kw))
  plain = _solve(x, y, a, b, "fp32", dict(epsilon=0.1), **dict(kw))
  np.testing.assert_array_equal(out.f.cpu().numpy(), plain.f.cpu().numpy())
  np.testing.assert_array_equal(out.g.cpu().numpy(), plain.g.cpu().numpy())
  assert out.reg_ot_cost == plain.reg_ot_cost


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
@pytest.mark.parametrize("taus,inner,momentum,zeros", [((0.8, 0.9), 1, 1.0, False),
                                                       ((0.95, 0.95), 10, 1.0, True),
                                                       ((0.7, 0.85), 5, 0.8, False)])
def test_recenter_potentials_unbalanced_both_sides(precision, taus, inner, momentum, zeros):
  """recenter_potentials=True with tau_a, tau_b < 1: the translation-invariant
  iteration of lse_step (sinkhorn.py:803-857), errors of the recentred state
  (sinkhorn.py:57-68) and recentred outputs (sinkhorn.py:1008-1009), against
  the oracle; the outputs are translation-optimal (a zero Sejourne shift)."""
  from ott_amd.solvers.linear import acceleration
  ta, tb = taus
  x, y, a, b = _cloud(211, 157, 5, seed=23, zero_a=[4] if zeros else None, zero_b=[9] if zeros else None)
  b = (b * 1.3).astype(np.float32)
  kw = dict(inner_iterations=inner, threshold=1e-4, max_iterations=400)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.1), recenter_potentials=True, tau_a=ta, tau_b=tb,
               momentum=acceleration.Momentum(value=momentum), **kw)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.1, recenter_potentials=True, tau_a=ta, tau_b=tb,
                                momentum=momentum, **kw)
  assert ref.converged
  _check_solve(out, ref, precision)
  f = out.f.cpu().numpy().astype(np.float64)
  g = out.g.cpu().numpy().astype(np.float64)
  if zeros:
    assert np.isneginf(f[4]) and np.isneginf(g[9])
  fr, _ = oracle.recenter(f, g, a.astype(np.float64), b.astype(np.float64), 0.1, ta, tb)
  fin = np.isfinite(f)
  assert np.max(np.abs(fr[fin] - f[fin])) <= 1e-4 * np.abs(f[fin]).max()


@pytest.mark.parametrize("mode", ["mean", "max_norm", "max_bound", "max_cost"])
def test_scale_cost_modes(mode):
  """PointCloud(scale_cost=...) string modes (pointcloud.py:342-383): the
  inverse scale against numpy in float64, then a solve on the scaled cost
  against the oracle given the same scale as a float."""
  x, y, a, b = _cloud(333, 257, 6, seed=17, shift=0.5)
  geom = _geom(x, y, "fp32", epsilon=0.05, relative_epsilon="mean", scale_cost=mode)
  x64, y64 = x.astype(np.float64), y.astype(np.float64)
  nx, ny = (x64 ** 2).sum(1).max(), (y64 ** 2).sum(1).max()
  ref = {
      "mean": ((x64[:, None, :] - y64[None]) ** 2).sum(-1).mean(),
      "max_norm": max(nx, ny),
      "max_bound": nx + ny + 2.0 * np.sqrt(nx * ny),
      "max_cost": ((x64[:, None, :] - y64[None]) ** 2).sum(-1).max(),
  }[mode]
  assert math.isclose(1.0 / geom.inv_scale_cost, ref, rel_tol=1e-5), (1.0 / geom.inv_scale_cost, ref)
  out = _solve(x, y, a, b, "fp32", dict(epsilon=0.05, relative_epsilon="mean", scale_cost=mode))
  ref_out = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, relative_epsilon="mean",
                                    scale_cost=1.0 / geom.inv_scale_cost)
  _check_solve(out, ref_out, "fp32")


def _median_f32(x, y):
  """np.median of the float32 costs summed from the differences (the entries
  the kernel selects from); float64 where the inputs make the sums exact."""
  c = ((x[:, None, :].astype(np.float64) - y[None].astype(np.float64)) ** 2).sum(-1)
  return float(np.median(c))


@pytest.mark.parametrize("n,m,d,kind", [(333, 257, 6, "normal"), (300, 258, 6, "normal"), (1, 1, 3, "normal"),
                                        (1, 2, 3, "normal"), (65, 64, 130, "normal"), (2048, 3001, 17, "normal"),
                                        (97, 96, 3, "integer"), (120, 80, 2, "duplicates")])
def test_median_cost_matches_numpy(n, m, d, kind):
  """ottx_cost_median (jnp.median of the unscaled cost matrix) against
  np.median of every entry: odd and even counts, one entry, d above the
  staging chunk, ties (integer coordinates give exact, repeated costs;
  duplicated points give repeated zeros)."""
  rng = np.random.default_rng(n + m + d)
  if kind == "integer":
    x = rng.integers(-3, 4, size=(n, d)).astype(np.float32)
    y = rng.integers(-3, 4, size=(m, d)).astype(np.float32)
  elif kind == "duplicates":
    base = rng.normal(size=(10, d)).astype(np.float32)
    x = base[rng.integers(0, 10, size=n)]
    y = base[rng.integers(0, 10, size=m)]
  else:
    x = rng.normal(size=(n, d)).astype(np.float32)
    y = (rng.normal(size=(m, d)) + 0.3).astype(np.float32)
  print(f"median n={n} m={m} d={d} {kind}", flush=True)
  geom = _geom(x, y, "fp32", epsilon=0.1, scale_cost="median")
  want = _median_f32(x, y)
  got = 1.0 / geom.inv_scale_cost
  if kind == "integer":
    assert got == want, (got, want)
  else:
    assert math.isclose(got, want, rel_tol=2e-6, abs_tol=1e-30), (got, want)
  assert math.isclose(geom.median_cost_matrix, 1.0, rel_tol=1e-12) or want == 0.0


def test_median_cost_nan_entry():
  x, y, _, _ = _cloud(50, 40, 3, seed=2)
  x[7, 1] = np.nan
  assert math.isnan(_geom(x, y, "fp32", epsilon=0.1, scale_cost=2.0).median_cost_matrix)


@pytest.mark.parametrize("precision", ["fp32", "bf16_split"])
def test_solve_with_median_scale(precision):
  """scale_cost="median" through the solve against the oracle given the
  oracle's own median as the scale."""
  x, y, a, b = _cloud(333, 257, 6, seed=17, shift=0.5)
  out = _solve(x, y, a, b, precision, dict(epsilon=0.05, relative_epsilon="mean", scale_cost="median"))
  med = oracle.median_cost(x, y)
  ref = oracle.solve_pointcloud(x, y, a, b, epsilon=0.05, relative_epsilon="mean", scale_cost=med)
  _check_solve(out, ref, precision)
